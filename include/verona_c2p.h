/*
 * verona_c2p.h -- batched conserved-to-primitive recovery (con2prim) with a
 * learned Newton initial guess, part of libverona_hip.so (gfx950).
 *
 * Reference counterpart: UtoP(U, P, eos, n_iter, tol), the con2prim kernel
 * object of Verona3D (src/3dim/structs.jl:218-628, called at
 * src/3dim/algos.jl:290 and :318).  Its first and main method is the 1DW
 * Newton iteration on W = w Gamma^2 (src/3dim/structs.jl:273-343) started
 * from the caller's previous primitives P (in-out argument = warm start).
 * The batch entry points below run that same iteration on a flat list of
 * cells with a choice of initial guess:
 *
 *   VRN_GUESS_GENERIC   P = (1, 1, 0, 0, 0) for every cell (no history)
 *   VRN_GUESS_WARM      the caller's previous primitives (UtoP's in-out P)
 *   VRN_GUESS_ANALYTIC  W0 = E = -T^t_t, a lower bound of W
 *   VRN_GUESS_LEARNED   W0 from a 4 -> 16 -> 16 -> 2 MLP evaluated with BF16
 *                       MFMA (FP32 accumulation) inside the solve kernel
 *
 * Every guess is finished by the same FP64 Newton iteration (the reference's
 * W_min clamp, v^2 clamp and halving), stopped when |dW| < rtol * W, with no
 * Lorentz-factor cap; a cell that does not converge falls back to the rest of
 * the reference cascade (2D Newton, 5D Newton, bisection;
 * src/3dim/structs.jl:346-596) and a cell that still fails returns NaN.
 *
 * Conventions as in verona_hip.h: 0 = ok, negative = error with the message
 * in vrn_c2p_last_error(); host pointers are borrowed for the call only; a
 * handle owns its device buffers and one HIP stream and is not re-entrant.
 * Cell arrays use Verona's variable-fastest order: U[5*i + v] with
 * v = (D, T^t_t, T^t_x, T^t_y, T^t_z) and P[5*i + v] with
 * v = (rho, u, u^x, u^y, u^z).
 */
#ifndef VERONA_C2P_H
#define VERONA_C2P_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VRN_GUESS_GENERIC 0
#define VRN_GUESS_WARM 1
#define VRN_GUESS_ANALYTIC 2
#define VRN_GUESS_LEARNED 3

/* MLP 4 -> 16 -> 16 -> 2, parameters packed as W1[16][4], b1[16], W2[16][16],
 * b2[16], W3[2][16], b3[2] (row = output unit), FP32. */
#define VRN_MLP_NIN 4
#define VRN_MLP_NHID 16
#define VRN_MLP_NOUT 2
#define VRN_MLP_NPARAM 386

typedef struct vrn_c2p vrn_c2p;

typedef struct vrn_c2p_stats {
    int64_t n;              /* cells solved */
    int64_t n_unrecovered;  /* cells whose P is not finite after the cascade */
    int64_t n_fallback;     /* cells the 1DW Newton did not finish */
    int64_t n_guess_bad;    /* learned guesses replaced by W0 = E (unphysical features) */
    int64_t iters_sum;      /* Newton iterations summed over cells (1DW + fallback) */
    int64_t iters_max;
    int64_t wave_max_sum;   /* per 64-cell wavefront: max iterations, summed */
    int64_t n_waves;
    int64_t hist[64];       /* cells per iteration count (last bin: >= 63) */
    double ms;              /* solve kernel time, HIP events on the handle's stream */
} vrn_c2p_stats;

/* Handle for batches of up to n_max cells, ideal-gas index gamma.
 * No counterpart in the reference (its UtoP works on the grid arrays). */
int vrn_c2p_create(int64_t n_max, double gamma, vrn_c2p **out);
int vrn_c2p_destroy(vrn_c2p *h);
const char *vrn_c2p_last_error(void);

/* Copy n cells of conserved variables (and optionally the previous
 * primitives, used by VRN_GUESS_WARM) to the device.  P_prev may be NULL. */
int vrn_c2p_upload(vrn_c2p *h, int64_t n, const double *U, const double *P_prev);

/* Recover P for the uploaded cells (the UtoP of src/3dim/structs.jl:218-628
 * on a flat batch): guess = VRN_GUESS_*, at most n_iter 1DW iterations,
 * relative stopping tolerance rtol.  stats may be NULL. */
int vrn_c2p_solve(vrn_c2p *h, int guess, int n_iter, double rtol, vrn_c2p_stats *stats);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the solve of vrn_c2p_solve (same kernels, same results
 * bit for bit) on caller-owned device arrays, with no handle and no
 * allocation.  U, P_prev (VRN_GUESS_WARM only, else may be NULL) and P are
 * device pointers to 5 planes of n doubles each (structure of arrays: variable
 * v of cell i at ptr[v * n + i]).  guess is VRN_GUESS_GENERIC, _WARM or
 * _ANALYTIC (the learned guess needs a handle's network).  Work is enqueued
 * on `stream` (a hipStream_t; NULL = the null stream) and the call returns;
 * with stats != NULL it waits for the stream and fills them (ms = 0: not
 * timed).  The working memory is a fixed device array inside the library,
 * sized for n <= VRN_C2P_NOALLOC_MAX_CELLS (about 19 MB on every GPU the
 * library runs on, held until the process ends); larger n is rejected.  All
 * calls on one GPU share that array: calls that overlap on one GPU (other
 * streams or host threads) race and must be serialised by the caller. */
#define VRN_C2P_NOALLOC_MAX_CELLS (1 << 20)
int vrn_c2p_solve_noalloc(int64_t n, const double *U, const double *P_prev, double *P, double gamma,
                          int guess, int n_iter, double rtol, void *stream, vrn_c2p_stats *stats);

/* Results of the last solve: primitives [n][5], the W each cell started
 * from, and its iteration count (capped at 255).  Any pointer may be NULL. */
int vrn_c2p_download(vrn_c2p *h, double *P, double *W0, uint8_t *iters);

/* Network parameters (VRN_MLP_NPARAM floats) used by VRN_GUESS_LEARNED; the
 * kernel uses their BF16 rounding. */
int vrn_mlp_set(vrn_c2p *h, const float *params);
int vrn_mlp_get(vrn_c2p *h, float *params);

/* Features X[n][VRN_MLP_NIN] and training targets T[n][2] of (U, P) pairs as the
 * learned guess defines them; valid[i] = 0 marks pairs outside the model's
 * domain.  P may be NULL (then T is not written). */
int vrn_mlp_features(vrn_c2p *h, int64_t n, const double *U, const double *P, float *X, float *T,
                     uint8_t *valid);

/* Network outputs Y[n][2] for features X[n][VRN_MLP_NIN]: the BF16 MFMA inference
 * kernel of the solve, on its own. */
int vrn_mlp_infer(vrn_c2p *h, int64_t n, const float *X, float *Y);

/* Loss (mean over samples of s (y_0 - t_0)^2 + w_1 (y_1 - t_1)^2) and its
 * gradient with respect to params, by the training kernels.  qat = 1
 * evaluates the network the way the inference kernel does (BF16 rounding of
 * the inputs, W1, W2, W3 and both hidden layers; FP32 b1, b2, b3 and
 * accumulation; straight-through gradients).  s = 1 + (q / q_ref)^2 with
 * q = X_2^2 e^{t_0} (= log(W/E) of the pair: the guess's relative W error is
 * q times the error of y_0) when q_ref > 0, else 1. */
int vrn_mlp_grad(vrn_c2p *h, int64_t n, const float *X, const float *T, const float *params,
                 int qat, float q_ref, float *loss, float *grad);

typedef struct vrn_mlp_train_params {
    int steps;        /* Adam steps */
    int batch;        /* samples per step (<= 65536) */
    float lr;         /* peak learning rate (cosine decay to lr_min) */
    float lr_min;
    float beta1, beta2, eps;
    float w_out1;     /* loss weight of the second output (log Gamma) */
    float q_ref;      /* > 0: weight of the first output's error, as in vrn_mlp_grad */
    int qat;          /* train through the inference kernel's BF16 rounding */
    uint64_t seed;    /* initialisation and minibatch sampling */
} vrn_mlp_train_params;

/* Train the network on the GPU from n (U, P) pairs (the solver's own
 * dumps and synthetic states): features and targets on the device,
 * minibatch forward/backward kernels, deterministic gradient reduction,
 * Adam.  If init is non-NULL training starts from those parameters, else
 * from a seeded He initialisation.  loss_curve (steps floats) may be NULL.
 * The trained parameters become the ones VRN_GUESS_LEARNED uses. */
int vrn_mlp_train(vrn_c2p *h, int64_t n, const double *U, const double *P, const float *init,
                  const vrn_mlp_train_params *tp, float *loss_curve, double *ms);

#ifdef __cplusplus
}
#endif

#endif
