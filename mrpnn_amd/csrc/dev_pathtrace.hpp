// Delta-tracking path sample shared by the judge (pathtrace.hip) and the interactive
// progressive renderer (interactive.hip): CalculateRadiance (core/render.cuh:535-597) with
// Woodcock tracking DeterminateNextVertex (core/render.cuh:399-434) and next-event estimation
// through the stochastic Tr (core/render.cuh:60-103), the TR mip 0 of Fill_TR, or the cached
// stochastic bake of Fill_BakeTR (render.cuh:589 reads whichever TR mip 0 holds); and the
// point-radiance paths of GetSample (core/render.cuh:615-654).
#pragma once
#include "dev_common.hpp"

namespace mrp {

__device__ inline float tr_stochastic(const DevVolume& v, PhiloxStream& rs, v3 pos, v3 dir, float dis, float alpha,
                                      float tr_scale, unsigned& taps) {
    const float smax = v.max_density * alpha;
    float tr = 1.0f, t = 0.0f;
    int loop = 0;
    while (loop++ < 10000 && tr > 0.5f) {
        float rk = rs.next();
        t -= logf(1.0f - rk) / smax * tr_scale;
        if (t > dis) break;
        float d = density_full(v, pos + (dir * t));
        ++taps;
        float S = d * alpha;
        tr *= 1.0f - fmaxf(0.0f, S / smax);
        if (d < 0.0f) t -= d;
    }
    // the reference's tr > 0.001 in double: float(0.001) is the smallest float above 0.001
    while (loop++ < 10000 && tr >= 0.001f) {
        float rk = rs.next();
        t -= logf(1.0f - rk) / smax * tr_scale;
        if (t > dis) break;
        float d = density_full(v, pos + (dir * t));
        ++taps;
        float S = d * alpha;
        if (rs.next() < fmaxf(0.0f, S / smax)) {
            tr = 0.0f;
            break;
        }
        if (d < 0.0f) t -= d;
    }
    return tr;
}

__device__ inline bool next_vertex(const DevVolume& v, PhiloxStream& rs, float alpha, float g, v3 pos, v3 dir,
                                   float dis, v3* np, v3* nd, unsigned& taps) {
    const float smax = v.max_density * alpha;
    float t = 0.0f;
    int loop = 0;
    while (loop++ < 10000) {
        float rk = rs.next();
        t -= logf(1.0f - rk) / smax;
        if (t > dis) return false;
        rk = rs.next();
        float d = density_full(v, pos + (dir * t));
        ++taps;
        float S = d * alpha;
        if (S / smax > rk) break;
        if (d < 0.0f) t -= d;
    }
    float e0 = rs.next();
    float e1 = rs.next();
    *nd = sample_hg(e0, e1, dir, g);
    *np = (dir * t) + pos;
    return true;
}

// Next-event transmittance source of the path tracer.
enum NeeMode : int {
    NEE_STOCHASTIC = 0,  // Tr of core/render.cuh:60-103 per vertex (unbiased)
    NEE_TR_MIP0 = 1,     // TR mip 0 of Fill_TR, 128^3 (render.cuh:589 with Update_TR's mips)
    NEE_CACHED = 2,      // the cached stochastic bake of Fill_BakeTR (render.cuh:589 after Update_CachedTR)
};
struct NeeSrc {
    int mode;
    const float* ctr;  // NEE_CACHED: ctr_res^3 texels, (x*res+y)*res+z
    int ctr_res;
};

__device__ inline float nee_tr(const DevVolume& vol, const NeeSrc& nee, PhiloxStream& rs, v3 pos, v3 light,
                               float alpha, float tr_scale, unsigned& taps) {
    if (nee.mode == NEE_TR_MIP0) return mip_tr(vol, 0, pos);
    if (nee.mode == NEE_CACHED) return sample_clamp_exact(nee.ctr, nee.ctr_res, pos + 0.5f);
    return tr_stochastic(vol, rs, pos, light, ray_box_distance(pos, light), alpha, tr_scale, taps);
}

// One path of pixel (row, col), sample s: its own Philox stream {seed, frame} /
// {pixel, RNG_PT, s, *}; the jitter comes first from that stream when vw.jitter < 0.
__device__ inline v3 pt_sample(const DevVolume& vol, const ViewDev& vw, v3 light, v3 fwd, float sx, float sy, int row,
                               int col, uint32_t pix, int s, int max_bounce, const NeeSrc& nee, float tr_scale,
                               unsigned& taps) {
    PhiloxStream rs(vw.seed, vw.frame, pix, RNG_PT, (uint32_t)s);
    float xu = vw.jitter, xv = vw.jitter;
    if (vw.jitter < 0.0f) {
        xu = rs.next();
        xv = rs.next();
    }
    float u = 1.0f - ((float)col + xu) / (float)vw.width;
    float vv = 1.0f - ((float)row + xv) / (float)vw.height;
    v3 dir = fwd + (vw.right * ((u * 2.0f - 1.0f) * sx)) + (vw.up * ((vv * 2.0f - 1.0f) * sy));
    dir = normalize3(dir);  // volume.cu:216
    dir = normalize3(dir);  // render.cuh:540
    v3 res = mk3(0, 0, 0);
    float dis = ray_box_offset(vw.ori, dir);
    if (dis < 0.0f) return sky_box(vol, dir);  // render.cuh:545-546
    {
        v3 spos = vw.ori + dir * dis, rdir = dir, csr = mk3(1, 1, 1);
        for (int b = 0; b < max_bounce; ++b) {
            v3 npos, ndir;
            float maxd = ray_box_distance(spos, rdir);
            if (!next_vertex(vol, rs, vw.alpha, vw.g, spos, rdir, maxd, &npos, &ndir, taps)) {
                res = res + sky_box(vol, rdir) * csr;  // render.cuh:569-573
                break;
            }
            spos = npos;
            csr = csr * vw.scatter_rate;
            float phase = henyey_greenstein(dot3(rdir, light), vw.g);
            float tr = nee_tr(vol, nee, rs, spos, light, vw.alpha, tr_scale, taps);
            res = res + vw.light_color * csr * (tr * phase);
            rdir = ndir;
        }
    }
    return res;
}

// GetSample (core/render.cuh:615-654): one path from point ori inside the medium.  The first
// direction is drawn from the HG lobe around dir, then delta tracking with next-event
// estimation through the stochastic Tr; each vertex scales the throughput by 1.001 * scatter
// (callers pass scatter_rate / 1.001), escapes add the sky.
__device__ inline v3 point_sample(const DevVolume& vol, PhiloxStream& rs, v3 ori, v3 dir, v3 light, v3 light_color,
                                  float scatter, float alpha, float g, int max_bounce, float tr_scale, unsigned& taps) {
    v3 res = mk3(0, 0, 0);
    const float e0 = rs.next();
    const float e1 = rs.next();
    v3 spos = ori, rdir = sample_hg(e0, e1, dir, g), csr = mk3(1, 1, 1);
    for (int b = 0; b < max_bounce; ++b) {
        v3 npos, ndir;
        const float dis = ray_box_distance(spos, rdir);
        if (!next_vertex(vol, rs, alpha, g, spos, rdir, dis, &npos, &ndir, taps)) {
            res = res + sky_box(vol, rdir) * csr;
            break;
        }
        spos = npos;
        csr = (csr * 1.001f) * scatter;
        const float ld = ray_box_distance(spos, light);
        const float phase = henyey_greenstein(dot3(rdir, light), g);
        res = res + light_color * csr * (tr_stochastic(vol, rs, spos, light, ld, alpha, tr_scale, taps) * phase);
        rdir = ndir;
    }
    return res;
}

}  // namespace mrp
