"""Second, independent restatement of RadiancePredictInternal (core/radiancePredict.cu:301-422)
in vectorised numpy float64: pixels are rows, layers are matrix products.  Used only to pin
the oracle's dataflow (block order, gate inputs, ADD3/REP3 aliasing, pooling slots, tail
scaling); it shares no code with the C++ oracle or the library.
"""
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASSETS = os.path.join(ROOT, "mrpnn_amd", "assets")


def load_weights():
    man = json.load(open(os.path.join(ASSETS, "mrpnn_weights.json")))
    blob = np.fromfile(os.path.join(ASSETS, "mrpnn_weights.bin"), "<f4").astype(np.float64)
    return {a["name"]: blob[a["offset"]:a["offset"] + a["count"]] for a in man["arrays"]}


_AMAX = None   # while predict(..., amax=d) runs: weight name -> largest |input| of that product
_NAMES = None  # id(weight array) -> weight name


def _lin(W, x, n_out):  # W[out][in] row-major; x: [P, in]
    if _AMAX is not None and x.size:
        nm = _NAMES[id(W)]
        _AMAX[nm] = max(_AMAX.get(nm, 0.0), float(np.abs(x).max()))
    return x @ W.reshape(n_out, -1).T


def _relu(x):
    return np.maximum(x, 0.0)


def _sig(x):
    return 1.0 / (1.0 + np.exp(-x))


def predict(w, feats, g, scatterrate, amax=None):
    """feats: [P, 577] rows [X_Val | X_Val_Hg | X_Val_Sub | gamma] -> [P, 3].  amax (a dict), when
    given, receives per weight matrix name the largest |input| of its product."""
    global _AMAX, _NAMES
    if amax is None:
        return _predict(w, feats, g, scatterrate)
    _AMAX, _NAMES = amax, {id(v): k for k, v in w.items()}
    try:
        return _predict(w, feats, g, scatterrate)
    finally:
        _AMAX = _NAMES = None


def _predict(w, feats, g, scatterrate):
    feats = np.asarray(feats, np.float64)
    P = feats.shape[0]
    X, H, S = feats[:, 0:192], feats[:, 192:384], feats[:, 384:576]
    gamma = feats[:, 576]
    src = [X, S, H]  # branch order A (val), B (sub/tr), C (hg)
    pool = np.zeros((P, 75))
    pool[:, 72] = g
    pool[:, 73] = gamma
    A2 = [np.zeros((P, 160)) for _ in range(3)]

    def block(last, frm, to, n, pid, sew0, sew1, names):
        for b in range(3):
            sl = src[b][:, frm:frm + n]
            pool[:, pid + b] = sl.mean(axis=1)
            pool[:, 36 + pid + b] = np.maximum(sl.max(axis=1), -10000.0)
        t = np.concatenate([pool[:, pid:pid + 3], pool[:, 36 + pid:36 + pid + 3], np.full((P, 1), g),
                            gamma[:, None]], axis=1)
        h = _relu(_lin(w[sew0], t, 8))
        s = _sig(_lin(w[sew1], h, 3))
        for b in range(3):
            a = src[b][:, frm:frm + n] * s[:, b:b + 1]
            if last is not None:
                a = a + A2[b][:, last:last + n]
            A2[b][:, to:to + n] = _relu(_lin(w[names[b] + "W"], a, n) + w[names[b] + "B"])

    def add3(frm, to, n):
        for b in range(3):
            A2[b][:, to:to + n] += src[b][:, frm:frm + n]

    def rep3(frm, to, n):
        for b in range(3):
            A2[b][:, frm:frm + n] = A2[b][:, to:to + n].copy()
            A2[b][:, to:to + n] = src[b][:, to:to + n]

    br = lambda k: ["L%d1" % k, "L_Tr%d1" % k, "L_Hg%d1" % k]
    dbr = lambda k: ["LD%d1" % k, "LD_Tr%d1" % k, "LD_Hg%d1" % k]
    block(None, 0, 0, 8, 0, "LSE01W", "LSE02W", br(0)); add3(0, 0, 8)
    block(0, 8, 8, 8, 3, "LSE11W", "LSE12W", br(1)); rep3(0, 8, 8)
    block(0, 16, 16, 16, 6, "LSE21W", "LSE22W", br(2)); add3(16, 16, 16)
    block(16, 32, 32, 16, 9, "LSE31W", "LSE32W", br(3)); add3(32, 32, 16)
    block(32, 48, 48, 16, 12, "LSE41W", "LSE42W", br(4)); rep3(32, 48, 16)
    block(32, 64, 64, 32, 15, "LSE51W", "LSE52W", br(5)); add3(64, 64, 32)
    block(64, 96, 96, 32, 18, "LSE61W", "LSE62W", br(6)); add3(96, 96, 32)
    block(96, 128, 128, 32, 21, "LSE71W", "LSE72W", br(7)); add3(128, 128, 32)
    comb = np.zeros((P, 128))
    for b, nm in enumerate(br(8)):
        comb[:, 32 * b:32 * b + 32] = _relu(_lin(w[nm + "W"], A2[b][:, 128:160], 32) + w[nm + "B"])
    block(None, 160, 0, 8, 24, "LDSE01W", "LDSE02W", dbr(0)); add3(160, 0, 8)
    block(0, 168, 8, 8, 27, "LDSE11W", "LDSE12W", dbr(1)); add3(168, 8, 8)
    block(8, 176, 16, 8, 30, "LDSE21W", "LDSE22W", dbr(2)); add3(176, 16, 8)
    block(16, 184, 24, 8, 33, "LDSE31W", "LDSE32W", dbr(3)); add3(184, 24, 8)
    for b, nm in enumerate(dbr(4)):
        comb[:, 96 + 8 * b:104 + 8 * b] = _relu(_lin(w[nm + "W"], A2[b][:, 24:32], 8) + w[nm + "B"])
    sr = np.asarray(scatterrate, np.float64)
    srp = sr ** 4

    def tail(srp_p):
        pl = pool.copy()
        pl[:, 74] = srp_p
        a = comb.copy()
        a[:, 120:128] = _relu(_lin(w["LGGSW"], pl[:, 72:75], 8) + w["LGGSB"])
        w6 = _sig(_lin(w["LSEFin2W"], _relu(_lin(w["LSEFin1W"], pl, 16)), 6))
        for k, (o, n) in enumerate([(0, 32), (32, 32), (64, 32), (96, 8), (104, 8), (112, 8)]):
            a[:, o:o + n] *= w6[:, k:k + 1]
        x = _relu(_lin(w["LC0W"], a, 128) + w["LC0B"])
        x = _relu(_lin(w["LC1W"], x, 64) + w["LC1B"])
        x = _relu(_lin(w["LC2W"], x, 32) + w["LC2B"])
        x = _relu(_lin(w["LXW"], x, 16) + w["LXB"])
        for k in range(6):
            y = _relu(_lin(w["LX%dW" % (2 * k)], x, 16) + w["LX%dB" % (2 * k)])
            x = x + _relu(_lin(w["LX%dW" % (2 * k + 1)], y, 16) + w["LX%dB" % (2 * k + 1)])
        return _relu(_lin(w["LX12W"], x, 1) + w["LX12B"])[:, 0]

    P0 = tail(srp[0])
    P1 = P0 if sr[1] == sr[0] else tail(srp[1])
    P2 = P0 if sr[2] == sr[0] else (P1 if sr[2] == sr[1] else tail(srp[2]))
    out = np.stack([P0, P1, P2], axis=1)
    return np.maximum(np.exp(out) - 1.0, 0.0) * srp[None, :]
