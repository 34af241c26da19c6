/* Learned gradient preconditioner: kernels of a 3D convolutional
 * encoder-decoder (inference and training) on MI355X (gfx950).
 *
 * No counterpart in the reference: the network stands in for the scaled
 * identity H0 = (y.s)/(y.y) of the coupled L-BFGS recursion
 * (step_4_3D_Forward_and_Inverse/LBFGS_Coupled.F90:95-104), which stays the
 * fallback.  The inversion driver calls these through its host mirror
 * (hirensi_amd/precond.py).
 *
 * Conventions: every pointer is a device pointer; `stream` is a hipStream_t
 * (NULL: the null stream) and every call only queues work on it; the return
 * value is 0 or a nonzero error code (argument errors: 1).  Volumes are NB
 * samples x C channels of (D, H, W), x fastest, FP32.  Convolution weights are
 * BF16 bit patterns [cout][cin * 27] (3x3x3 taps t = 9 (dz+1) + 3 (dy+1) +
 * (dx+1), zero padding); products are formed from BF16-rounded operands and
 * accumulated in FP32 on the matrix cores.
 */
#ifndef HIRENSI_NN_H
#define HIRENSI_NN_H

#ifdef __cplusplus
extern "C" {
#endif

/* out[b][co] = bias[co] + sum_{ci,t} w[co][ci,t] in[b][ci](v + offset(t))
 * (+ out when accumulate != 0).  bias may be NULL. */
int hr_nn_conv3d(const float* in, const unsigned short* w, const float* bias, float* out, int cin, int cout, int NB,
                 int D, int H, int W, int accumulate, void* stream);

/* Weight and bias gradients of hr_nn_conv3d for the output gradient dy:
 * dw[co][ci,t] = sum over samples and voxels of dy * in(shifted); db[co] = sum dy
 * (db may be NULL).  part: nslice * cout * cin * 27 floats of scratch (the
 * column range is split into nslice slices, summed in slice order). */
int hr_nn_conv3d_wgrad(const float* in, const float* dy, float* part, int nslice, float* dw, float* db, int cin,
                       int cout, int NB, int D, int H, int W, void* stream);

/* wt[ci][co * 27 + 26 - t] = w[co][ci * 27 + t]: weights of the input-gradient
 * convolution (hr_nn_conv3d of dy with wt gives dL/din). */
int hr_nn_transpose_weights(const unsigned short* w, unsigned short* wt, int cin, int cout, void* stream);

/* x = x > 0 ? x : slope x (in place); its backward dy *= (y > 0 ? 1 : slope). */
int hr_nn_leaky_relu(float* x, long long n, float slope, void* stream);
int hr_nn_leaky_relu_grad(const float* y, float* dy, long long n, float slope, void* stream);

/* 2x2x2 average pooling of C volumes of even size (D, H, W) and its backward
 * (dx = dy / 8 on the 8 fine voxels, added to dx when accumulate != 0). */
int hr_nn_avgpool2(const float* in, float* out, long long C, int D, int H, int W, void* stream);
int hr_nn_avgpool2_grad(const float* dy, float* dx, long long C, int D, int H, int W, int accumulate, void* stream);

/* nearest 2x upsampling of C coarse volumes to (D, H, W) and its backward. */
int hr_nn_upsample2(const float* in, float* out, long long C, int D, int H, int W, void* stream);
int hr_nn_upsample2_grad(const float* dy, float* dx, long long C, int D, int H, int W, void* stream);

/* loss[0] = mean (y - t)^2 over n; grad (may be NULL) = 2 (y - t) / n;
 * part: 256 floats of scratch. */
int hr_nn_mse(const float* y, const float* t, long long n, float* grad, float* part, float* loss, void* stream);

/* Adam step (step >= 1, bias-corrected) on FP32 master weights w with
 * moments m, v; wb (may be NULL) receives the BF16 copy. */
int hr_nn_adam(float* w, const float* g, float* m, float* v, unsigned short* wb, long long n, float lr, float b1,
               float b2, float eps, int step, void* stream);

/* wb = BF16(w), round to nearest even. */
int hr_nn_to_bf16(const float* w, unsigned short* wb, long long n, void* stream);

#ifdef __cplusplus
}
#endif

#endif
