/*
 * pgfem_learn.h — C-ABI of the learned initial guess for the Newton/PCG solves.
 *
 * No counterpart in the reference: the reference starts every KSP/SNES solve
 * from the previous iterate (src/SolverNonlinearSNES.cpp:197-217,
 * src/SolverNonlinearKSP.cpp:55-159).  This interface adds a predictor in
 * front of those solves; plain PCG remains the solver and the judge of
 * convergence.
 *
 * The network is a small 3D U-Net on a 64^3 voxel grid of the unit cube:
 *   input  [64^3][8]  FP32 (rounded to bf16 on upload), channels:
 *          0 inclusion indicator, 1 damage (delta_max / delta_c),
 *          2..4 step residual (normalised), 5..7 previous load increment (normalised)
 *   output [64^3][3]  displacement correction (normalised)
 * Voxel (x, y, z) has index (z*64 + y)*64 + x.  Training and inference run on
 * the device with bf16 MFMA products and FP32 accumulation; weights are FP32.
 */
#ifndef PGFEM_LEARN_H
#define PGFEM_LEARN_H

#include <stdint.h>

#include "pgfem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pgf_net pgf_net;

PGF_API int pgf_net_create(uint64_t seed, pgf_net** out);
PGF_API void pgf_net_destroy(pgf_net* net);
PGF_API const char* pgf_net_last_error(const pgf_net* net);
/* parameters, flattened layer by layer: W[t][cin][cout] then b[cout] */
PGF_API int pgf_net_num_params(pgf_net* net, int64_t* n);
PGF_API int pgf_net_get_params(pgf_net* net, float* w);
PGF_API int pgf_net_set_params(pgf_net* net, const float* w);
/* y = net(x), host buffers x [64^3][8], y [64^3][3] */
PGF_API int pgf_net_forward(pgf_net* net, const float* x, float* y);
/* gradient of sum(dy . net(x)) with respect to every parameter, host buffers */
PGF_API int pgf_net_backward(pgf_net* net, const float* x, const float* dy, float* grads);
/* training set kept in device memory: input [64^3][8], target [64^3][3] */
PGF_API int pgf_net_add_sample(pgf_net* net, const float* x, const float* target);
PGF_API int pgf_net_num_samples(pgf_net* net, int32_t* n);
/* `epochs` passes of Adam (one step per sample) on the mean squared error;
 * loss[e] = mean loss of epoch e (may be NULL) */
PGF_API int pgf_net_train(pgf_net* net, int32_t epochs, float lr, float* loss);

/* ---- use inside the Newton step (pgf_newton_step with use_guess 2 or 3) ----
 * attach: box = {x0, y0, z0, x1, y1, z1} of the voxel grid, indicator [64^3]
 * (may be NULL), net may be NULL (features only).  With a state attached every
 * Newton step records its start features; use_guess 2 starts the first PCG
 * solve from the Galerkin (K-energy optimal) combination of the network's
 * three prolonged output channels and the previous load increment, 3 from the
 * previous increment alone.  The combination never has a larger energy-norm
 * error than a zero start. */
PGF_API int pgf_learn_attach(pgf_ctx* ctx, pgf_net* net, const double* box, const float* indicator);
PGF_API int pgf_learn_detach(pgf_ctx* ctx);
/* after an accepted load step: its converged nodal field becomes the previous one */
PGF_API int pgf_learn_commit(pgf_ctx* ctx);
/* add (start features, normalised converged increment) of the last step to the net's training set */
PGF_API int pgf_learn_record(pgf_ctx* ctx);
/* out[0..3] Galerkin coefficients of the last guess, out[4] columns used, out[5] its device time (ms) */
PGF_API int pgf_learn_last(pgf_ctx* ctx, double* out);
/* host copy of the last start features [64^3][8] */
PGF_API int pgf_learn_features(pgf_ctx* ctx, float* out);

#ifdef __cplusplus
}
#endif
#endif
