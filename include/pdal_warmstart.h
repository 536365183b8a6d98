/*
 * Learned warm start for families of LPs that share A (the fixed-A family
 * b = b0 + U theta_b, c = c0 + V theta_c): a multilayer perceptron maps the
 * parameters theta to the coefficients of the solution in a POD basis, and
 * the basis expansion gives the primal-dual start of a batched solve
 * (pdal_batch.h).  No counterpart in the reference, which solves every
 * instance from zero.
 *
 * The MLP runs on the GPU with hand-written gfx950 kernels: BF16 weights and
 * activations as GEMM inputs (round to nearest even), FP32 MFMA accumulation
 * (v_mfma_f32_16x16x16_bf16), FP32 master weights updated by Adam.  Layers are
 * dense, ReLU between hidden layers, none after the last; the loss is the mean
 * squared error over the batch and the outputs.
 *
 * All arrays are host memory, row-major, float32 unless stated (the one
 * device address is pdal_pod_expand_device's output); every entry point
 * returns 0 on success and a negative value on error.
 */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pdal_mlp pdal_mlp_t;

/* widths[0..n_layers]: input width, hidden widths, output width.  Weights
 * are drawn He-uniform (+-sqrt(6/fan_in)) from std::mt19937_64(seed), biases 0. */
pdal_mlp_t *pdal_mlp_create(int n_layers, const int *widths, unsigned long long seed);
void pdal_mlp_free(pdal_mlp_t *mlp);

/* Parameters in layer order, each layer W (out x in, row-major) then b (out). */
long long pdal_mlp_param_count(const pdal_mlp_t *mlp);
int pdal_mlp_get_params(const pdal_mlp_t *mlp, float *out);
int pdal_mlp_set_params(pdal_mlp_t *mlp, const float *in);

/* y (batch x out) = MLP(x (batch x in)). */
int pdal_mlp_forward(pdal_mlp_t *mlp, const float *x, int batch, float *y);

/* Loss 1/(batch*out) * sum (MLP(x) - y)^2 and its gradient with respect to
 * every parameter (grad: param_count floats, same order; may be NULL). */
int pdal_mlp_loss_grad(pdal_mlp_t *mlp, const float *x, const float *y, int batch, float *loss, float *grad);

/* Train on n samples (x: n x in, y: n x out) for `epochs` passes in
 * minibatches of `batch` (a seeded shuffle per epoch; the last partial
 * minibatch is dropped), Adam(lr, 0.9, 0.999, 1e-8).  loss_hist (epochs
 * floats, may be NULL) receives each epoch's mean minibatch loss. */
int pdal_mlp_train(pdal_mlp_t *mlp, const float *x, const float *y, int n, int batch, int epochs, float lr,
                   unsigned long long seed, float *loss_hist);

/* out (batch x dim, float64) = mean (dim) + coef (batch x k) . basis (k x dim),
 * FP32 products accumulated in FP32, stored as FP64 (a start point). */
int pdal_pod_expand(const float *basis, const float *mean, int dim, int k, const float *coef, int batch,
                    double *out);

/* A POD basis kept on the device: basis (k x dim) and mean (dim) uploaded
 * once, with optional bounds lo, hi (dim float64 each, may be NULL) that clip
 * every expanded entry.  NULL on error. */
typedef struct pdal_pod pdal_pod_t;
pdal_pod_t *pdal_pod_create(const float *basis, const float *mean, int dim, int k, const double *lo,
                            const double *hi);
void pdal_pod_free(pdal_pod_t *pod);

/* out (batch x dim float64, a DEVICE address on the current device, e.g. the
 * start buffer of a batch: pdal_cp_al_batch_start_buffers) = clip(mean +
 * coef . basis), coef (batch x k float32) in host memory.  out_capacity is
 * the number of doubles `out` holds; a call whose batch x dim exceeds it
 * writes nothing and returns -1.  Returns when the expansion has finished. */
int pdal_pod_expand_device(pdal_pod_t *pod, const float *coef, int batch, double *out, long long out_capacity);

/* pdal_pod_expand_device for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops that forbid hipMalloc): no
 * counterpart in the reference.  The same expansion, with the coefficients
 * staged in a fixed device array of the library (131,072 floats, 512 KiB per
 * GPU, held until the process ends) instead of a buffer the pod allocates.
 * batch x k must not exceed 131,072 (else -3 and a message on stderr, nothing
 * written).  Calls that overlap on one GPU share the array and race.  The
 * pod, out and coef are the caller's, as for pdal_pod_expand_device. */
int pdal_pod_expand_device_noalloc(pdal_pod_t *pod, const float *coef, int batch, double *out,
                                   long long out_capacity);

#ifdef __cplusplus
}
#endif
