/* libqmpaint: the learned excess-attenuation surrogate's C ABI.
 *
 * No counterpart in the reference (it has no learned part): these entry points train and run the
 * small network that stands in for the ground + diffraction terms of a path state (the per-band
 * A_ground + A_dif of physics/src/cnossos/diffraction.rs:64-196, 8 octave bands) on the device,
 * with BF16 MFMA products and FP32 accumulation. The exact painter stays the judge of its error
 * and the fallback outside its feature envelope.
 *
 * Network: 40 inputs -> 4 hidden layers of 256 (tanh-form GELU) -> 8 outputs. Parameters are one
 * flat FP32 vector of QM_MLP_PARAMS floats: per layer the weights [out][in] row-major, then the
 * biases [out]. Loss: mean squared error over rows and outputs. Optimiser: Adam.
 *
 * Conventions as qm_paint.h: 0 on success, else a message in `err`; device pointers (`d_*`) are
 * the caller's, on the library's device; `stream` is a hipStream_t (NULL: the default stream).
 */
#ifndef QM_SURROGATE_H
#define QM_SURROGATE_H

#include <stddef.h>
#include <stdint.h>

#include "qm_paint.h"

#define QM_MLP_INPUTS 40
#define QM_MLP_HIDDEN 256
#define QM_MLP_OUTPUTS 8
#define QM_MLP_LAYERS 5
#define QM_MLP_PARAMS 209928

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qm_mlp qm_mlp;

/* A network on the device from `params` (host, QM_MLP_PARAMS floats), with training buffers for
 * up to `max_batch` rows a step (0: inference only). */
int qm_mlp_new(qm_gpu* gpu, const float* params, size_t max_batch, qm_mlp** out, char* err, size_t errlen);
/* The current FP32 parameters (host, QM_MLP_PARAMS floats). */
int qm_mlp_params(qm_mlp* mlp, float* params, char* err, size_t errlen);
void qm_mlp_free(qm_mlp* mlp);

/* The loss of `rows` rows (`d_inputs` rows x 40, `d_targets` rows x 8, FP32) and its gradient
 * with respect to every parameter (`d_gradient`, QM_MLP_PARAMS floats, overwritten); `loss`
 * (host) receives the loss. rows <= max_batch. No update. */
int qm_mlp_gradient_device(qm_mlp* mlp, const float* d_inputs, const float* d_targets, size_t rows, float* d_gradient,
                           float* loss, void* stream, char* err, size_t errlen);
/* One Adam step (learning rate `lr`, betas 0.9/0.999, eps 1e-8) on `rows` rows. `loss` (host,
 * optional) receives the step's loss before the update (it synchronises the stream). */
int qm_mlp_train_step_device(qm_mlp* mlp, const float* d_inputs, const float* d_targets, size_t rows, float lr,
                             float* loss, void* stream, char* err, size_t errlen);
/* Outputs (`d_outputs`, rows x 8 FP32) of `rows` rows of inputs (rows x 40 FP32). */
int qm_mlp_infer_device(qm_mlp* mlp, const float* d_inputs, size_t rows, float* d_outputs, void* stream, char* err,
                        size_t errlen);

/* Training samples: per pair (receiver `d_pair_receiver[i]` of `d_receivers` as
 * qm_receivers_device makes them, source `d_pair_source[i]` in upload order) the network's two
 * input rows (calm, then favourable: 2 x QM_MLP_INPUTS floats at d_inputs + 80*i), the exact
 * path's A_ground + A_dif per band of each state as targets (2 x 8 floats at d_targets + 16*i) and
 * `d_in_envelope[i]` (1 when the path lies in the network's feature envelope and the exact path
 * did not fail, else 0). The inputs are those of quietmap_amd/csrc/qm_features.hpp. */
int qm_surrogate_samples_device(qm_gpu* gpu, const qm_square* square, const qm_receiver* d_receivers,
                                const uint32_t* d_pair_receiver, const uint32_t* d_pair_source, size_t n_pairs,
                                float* d_inputs, float* d_targets, uint32_t* d_in_envelope, void* stream, char* err,
                                size_t errlen);

/* The painted grid with the surrogate: qm_paint_receivers_device's contract (receivers on a
 * row-major nx x ny grid, every source within `cutoff_m`), with the network's boundary terms for
 * every path inside its envelope and the exact path for the others. `d_surrogate_paths`
 * (optional) receives per receiver the paths the network answered. */
int qm_surrogate_paint_receivers_device(qm_gpu* gpu, const qm_square* square, qm_mlp* mlp,
                                        const qm_receiver* d_receivers, uint32_t nx, uint32_t ny, double cutoff_m,
                                        double* d_period_energy, uint32_t* d_paths, uint32_t* d_failures,
                                        uint32_t* d_surrogate_paths, void* stream, char* err, size_t errlen);

#ifdef __cplusplus
}
#endif

#endif
