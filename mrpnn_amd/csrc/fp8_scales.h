// FP8 activation exponents of the MFMA network's 60 layers (mlp_mfma.h, MfmaLayerId order):
// the inputs of layer id are stored as e4m3(clamp(x, +-448 * 2^-e) * 2^e), e = kFp8ActExp[id],
// the largest e with (max |input|) * 2^e <= 448.  Calibrated by mrpnn_amd/assets/calibrate_fp8.py
// (BF16 network over 7761032 active-pixel feature rows of seeded fBm views: 512^3 clouds, alpha 64,
// 32 suns; 128^3-384^3 clouds, alpha 16-128; albedo 0.99 / 0.999 / distinct channels).
#pragma once

namespace mrp {
// per layer: max |input| seen in calibration, then the exponent
constexpr int kFp8ActExp[60] = {
       9,  // layer  0: max |input| 0.564108
       8,  // layer  1: max |input| 1.4645
       7,  // layer  2: max |input| 2.55582
       7,  // layer  3: max |input| 2.24347
       7,  // layer  4: max |input| 2.6103
       7,  // layer  5: max |input| 3.31264
       6,  // layer  6: max |input| 3.66104
       6,  // layer  7: max |input| 4.75546
       6,  // layer  8: max |input| 4.96278
       9,  // layer  9: max |input| 0.548824
       7,  // layer 10: max |input| 2.12856
       6,  // layer 11: max |input| 3.68373
       6,  // layer 12: max |input| 6.70609
       5,  // layer 13: max |input| 10.651
       5,  // layer 14: max |input| 9.37332
       4,  // layer 15: max |input| 16.0503
       4,  // layer 16: max |input| 23.1622
       3,  // layer 17: max |input| 28.2864
       8,  // layer 18: max |input| 1.27926
       7,  // layer 19: max |input| 3.08844
       8,  // layer 20: max |input| 1.63674
       7,  // layer 21: max |input| 2.64121
       7,  // layer 22: max |input| 2.42396
       7,  // layer 23: max |input| 3.43622
       7,  // layer 24: max |input| 2.30122
       7,  // layer 25: max |input| 1.83592
       7,  // layer 26: max |input| 2.2448
       9,  // layer 27: max |input| 0.643283
       8,  // layer 28: max |input| 1.62176
       5,  // layer 29: max |input| 7.15254
       4,  // layer 30: max |input| 16.5323
       9,  // layer 31: max |input| 0.727986
       7,  // layer 32: max |input| 2.99962
       5,  // layer 33: max |input| 7.74737
       5,  // layer 34: max |input| 9.0914
       8,  // layer 35: max |input| 1.17136
       5,  // layer 36: max |input| 8.0753
       5,  // layer 37: max |input| 9.43524
       5,  // layer 38: max |input| 11.9472
       5,  // layer 39: max |input| 11.6536
       5,  // layer 40: max |input| 9.31028
       6,  // layer 41: max |input| 6.66532
       6,  // layer 42: max |input| 5.9615
       5,  // layer 43: max |input| 10.762
       5,  // layer 44: max |input| 10.2442
       5,  // layer 45: max |input| 12.9853
       6,  // layer 46: max |input| 5.99625
       8,  // layer 47: max |input| 1.55131
      10,  // layer 48: max |input| 0.436263
       8,  // layer 49: max |input| 1.55131
       8,  // layer 50: max |input| 1.45518
       8,  // layer 51: max |input| 1.55131
       9,  // layer 52: max |input| 0.693425
       8,  // layer 53: max |input| 1.55131
      10,  // layer 54: max |input| 0.402145
       8,  // layer 55: max |input| 1.55131
       9,  // layer 56: max |input| 0.724329
       8,  // layer 57: max |input| 1.55131
      10,  // layer 58: max |input| 0.415219
       8,  // layer 59: max |input| 1.55131
};
}  // namespace mrp
