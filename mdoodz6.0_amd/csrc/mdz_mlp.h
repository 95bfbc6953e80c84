/*
 * mdz_mlp.h -- the surrogate network that supplies the initial guess of the
 * visco-elastic local iteration: MLP 9 -> 64 -> 64 -> 64 -> 2 with SiLU,
 * BF16 operands on the matrix cores, FP32 accumulation.
 *
 * Features (dimensional inputs, fixed affine maps to about [-1, 1]):
 *   f0 = (log10(Eii [1/s]) + 15) / 3       visco-elastic effective strain rate invariant
 *   f1 = (T [K] - 1123) / 850
 *   f2 = (P [GPa] - 5) / 5
 *   f3 = (tau0 [MPa] - 250) / 250           old stress invariant sqrt(txx0^2 + txz0^2)
 *   f4 = log10(dt [s]) - 11
 *   f5..f7 = one-hot phase 0, 1, 2
 *   f8 = (log10(tau_y [MPa]) - 1.5) / 1.5   Drucker-Prager yield stress C cos(phi) + P sin(phi)
 * Outputs:
 *   y0 = 4 log10(eta_ve / eta_ref)          root of the visco-elastic iteration relative to
 *                                           the reference's guess eta_ref (the smallest
 *                                           isolated viscosity, RheologyDensity.c:355)
 *   y1 = (log10(max(gdot [1/s], 1e-22)) + 16) / 6   plastic multiplier
 *
 * Parameter blob (float, MDZ_MLP_NPARAMS): W1[64][9] b1[64] W2[64][64] b2[64]
 * W3[64][64] b3[64] W4[2][64] b4[2], weights row-major [out][in].
 */
#pragma once

#include "mdz_device.h"

#define MLP_IN 9
#define MLP_H 64
#define MLP_OUT 2
#define MLP_OFF_W1 0
#define MLP_OFF_B1 (MLP_OFF_W1 + MLP_H * MLP_IN)
#define MLP_OFF_W2 (MLP_OFF_B1 + MLP_H)
#define MLP_OFF_B2 (MLP_OFF_W2 + MLP_H * MLP_H)
#define MLP_OFF_W3 (MLP_OFF_B2 + MLP_H)
#define MLP_OFF_B3 (MLP_OFF_W3 + MLP_H * MLP_H)
#define MLP_OFF_W4 (MLP_OFF_B3 + MLP_H)
#define MLP_OFF_B4 (MLP_OFF_W4 + MLP_OUT * MLP_H)
#define MLP_NPARAMS (MLP_OFF_B4 + MLP_OUT)

/* Padded BF16 weights for the matrix-core tiles: W1p[64][16] (inputs 9..15
 * zero), W2p[64][64], W3p[64][64], W4p[16][64] (outputs 2..15 zero), and the
 * transposes used by the backward pass. */
#define MLP_WP_W1 0
#define MLP_WP_W2 (MLP_WP_W1 + 64 * 16)
#define MLP_WP_W3 (MLP_WP_W2 + 64 * 64)
#define MLP_WP_W4 (MLP_WP_W3 + 64 * 64)
#define MLP_WP_W2T (MLP_WP_W4 + 16 * 64)
#define MLP_WP_W3T (MLP_WP_W2T + 64 * 64)
#define MLP_WP_W4T (MLP_WP_W3T + 64 * 64)
#define MLP_WP_TOTAL (MLP_WP_W4T + 64 * 16)
/* padded biases: b1[64] b2[64] b3[64] b4[16] */
#define MLP_BP_TOTAL (64 * 3 + 16)

/* Device copy of a network in kernel layout. */
struct DevMLPW {
    const unsigned short *wp; /* MLP_WP_TOTAL bf16 */
    const float *bp;          /* MLP_BP_TOTAL */
};

/* blob -> kernel layout (one launch) */
void mdz_mlp_pack(const float *blob, unsigned short *wp, float *bp, hipStream_t s);

/* Grid guess: factor on the reference's initial guess of eta_ve for the
 * dominant phase of each centroid / vertex, into guess_n[p*nc + c] /
 * guess_s[p*nv + v] (0 for the other phases, and outside the sampled
 * domain: those start from the reference guess).  Needs the vertex fields
 * of the rheology pass (T_s, P_s) in place. */
void mdz_launch_mlp_guess(const DevGrid &g, const DevParams &prm, DevMLPW w, float *guess_n, float *guess_s,
                          hipStream_t s);
