/*
 * nnx_host.h - host-side builders of libnnx_host.so (plain C++, no GPU).
 *
 * Network construction stays on the host, as in the reference
 * (Network::Generate, src/network_factory.cuh:23-38); these builders produce
 * the synthetic inputs of the benchmark configurations and the CSR/CSC
 * layouts the device kernels consume.  Every random draw comes from
 * Philox4x64-10 (Salmon et al., SC'11), one counter stream per
 * (seed, purpose, item), so results do not depend on the thread count.
 * All arrays are host memory owned by the caller.
 */
#ifndef NNX_HOST_H
#define NNX_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x64-10 block function: out = philox(ctr, key).                    */
void nnx_philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);

/* Streams.  Value i of stream (seed, purpose) is word i%4 of
 * philox(ctr = {i/4, purpose, 0, 0}, key = {seed, 0}).
 * uniform: (w >> 11) * 2^-53 in [0,1); normal: Box-Muller on two uniforms
 * (value i uses words 2i, 2i+1 of the stream, cos branch).                 */
void nnx_gen_uniform(uint64_t seed, uint64_t purpose, int64_t n, double lo, double hi, double* out);
void nnx_gen_normal(uint64_t seed, uint64_t purpose, int64_t n, double mu, double sigma, double* out);
void nnx_gen_lognormal(uint64_t seed, uint64_t purpose, int64_t n, double mu, double sigma, double* out);

/* Erdos-Renyi with fixed out-degree: k distinct targets per source drawn
 * uniformly (self-loops excluded unless allow_self), each source's targets
 * sorted ascending.  Same shape as UniformRandomNetwork
 * (src/network_factory.cu:96-131).  offsets: n+1, targets: n*k.
 * Returns 0, or -1 if k exceeds the number of candidates.                   */
int nnx_gen_er_fixed(uint64_t seed, int64_t n, int32_t k, int allow_self, int32_t* offsets,
                     int32_t* targets);

/* Uniform distinct targets with per-source degrees offsets[i+1]-offsets[i]
 * (same draw as nnx_gen_er_fixed, sorted per source).  Returns 0, or -1 if a
 * degree exceeds the number of candidates.                                 */
int nnx_gen_uniform_targets(uint64_t seed, int64_t n, const int32_t* offsets, int allow_self,
                            int32_t* targets);

/* Out-degrees k_i = max(1, round(LogNormal(mu, sigma))), capped at cap.
 * Fills outdeg (n) and returns the total.                                  */
int64_t nnx_gen_outdeg_lognormal(uint64_t seed, int64_t n, double mu, double sigma, int32_t cap,
                                 int32_t* outdeg);

/* Out-degrees k_i ~ Poisson(lambda), fills outdeg (n), returns the total.  */
int64_t nnx_gen_outdeg_poisson(uint64_t seed, int64_t n, double lambda, int32_t* outdeg);

/* Distance-dependent targets: for each source i, offsets[i+1]-offsets[i]
 * distinct targets j != i drawn without replacement with probability
 * proportional to exp(-|x_j - x_i| / lambda) over the neurons with
 * |x_j - x_i| <= cutoff (exact: proposal from per-cell upper bounds, then
 * acceptance).  pos: 3n floats (x,y,z interleaved) inside [0, box)^3.
 * Targets of each source sorted ascending.  Returns 0, or -1 if a source
 * has fewer candidates than its degree (its degree is then reduced and the
 * return is the number of reduced sources, negated).                        */
int64_t nnx_gen_spatial_targets(uint64_t seed, int64_t n, const float* pos, double box,
                                double lambda, double cutoff, int32_t* offsets, int32_t* targets);

/* CSR (out-edges) -> CSC (in-edges sorted by source, ties by CSR position),
 * out_to_in[e] = CSC position of CSR edge e; in_degree (optional, float)
 * is the reference's COUNT array (src/network_factory.cu:20-23).           */
void nnx_csr_to_csc(int64_t n, int64_t nnz, const int32_t* off, const int32_t* tgt,
                    const float* w, int32_t* in_off, int32_t* in_src, float* in_w,
                    int32_t* out_to_in, float* in_degree);

/* Steady-state gates m,h,n = alpha/(alpha+beta) at V (double), with the
 * reference's rate functions evaluated at V + v_shift.                      */
void nnx_hh_steady_state(int64_t n, const double* V, double v_shift, double* m, double* h,
                         double* ng);

/* voxel of each neuron: floor(pos/h) per axis, clamped to the grid.       */
void nnx_voxel_of(int64_t n, const float* pos, double h, int32_t nx, int32_t ny, int32_t nz,
                  int32_t* voxel);

/* Reference trilinear interpolation matrix (N x G CSR, 8 entries per row):
 * the grid spans [-(nx-1)h/2, (nx-1)h/2] per axis; base index and fractions
 * as in Scatter_grid_converter_3d::initialize_distribute_private
 * (src/paracrine.cu:136-274).  offsets n+1, cols/vals 8n.  Returns the
 * number of neurons whose 8 corners do not all lie inside the grid (their
 * out-of-grid corners are clamped onto the grid and given weight 0).       */
int64_t nnx_trilinear_matrix(int64_t n, const float* x, const float* y, const float* z,
                             int32_t nx, int32_t ny, int32_t nz, double h, int32_t* offsets,
                             int32_t* cols, float* vals);

/* Transpose of a general CSR (rows x cols) with values, rows of the result
 * listing their entries in ascending original-row order.                   */
void nnx_csr_transpose(int64_t rows, int64_t cols, const int32_t* off, const int32_t* idx,
                       const float* val, int32_t* t_off, int32_t* t_idx, float* t_val);

/* Gaussian random field, unit variance, covariance exp(-r^2 / (2 corr^2))
 * (lengths in voxels): white noise from the Philox stream (seed, purpose)
 * filtered by a separable Gaussian kernel.                                 */
void nnx_gen_grf(uint64_t seed, uint64_t purpose, int32_t nx, int32_t ny, int32_t nz, double corr,
                 float* out);

/* Face coefficients (nnx_grid kx/ky/kz layout) from per-voxel coefficients:
 * harmonic mean across interior faces, the voxel's value on boundary faces. */
void nnx_face_coefficients(int32_t nx, int32_t ny, int32_t nz, const float* c, float* kx, float* ky,
                           float* kz);

/* Electrode regions (Morphology::GenSphericalRegions, src/morphology.cpp:
 * 48-97, src_wholebrain/morphology.cpp): region r = the k neurons nearest to
 * query r, nearest first, by squared Euclidean distance computed in FP64
 * from the float positions (pos: 3n, x,y,z interleaved), ties by lower index.
 * Queries are neuron indices (center_idx, a neuron is its own nearest) or
 * points (center_pos: 3 n_regions floats; center_idx NULL).  idx: n_regions*k
 * (the reference's knn order, from its exact ANN search), offsets:
 * n_regions+1 = r*k.  The reference copies k+1 entries per region into k
 * slots (the last region writes one past the end); here exactly k.
 * Returns 0, or -1 for k outside [1, n] or a center index outside [0, n).  */
int nnx_spherical_regions(int64_t n, const float* pos, int64_t n_regions, const int32_t* center_idx,
                          const float* center_pos, int32_t k, int32_t* offsets, int32_t* idx);

/* ---- files ---------------------------------------------------------------
 * HDF5 datasets without the HDF5 library (a reader of the published file
 * format: superblocks 0-3, symbol-table and compact groups, integer and IEEE
 * float datasets of either byte order, compact / contiguous / chunked
 * storage with version-1 B-tree, single-chunk, implicit or fixed-array chunk
 * indexes, deflate / shuffle / fletcher32).  `name` is a path such as
 * "/row_ptr" or "grp/x".  Errors return -1 with a message in err.         */
enum { NNX_H5_I32 = 0, NNX_H5_F32 = 1, NNX_H5_F64 = 2, NNX_H5_I64 = 3 };
/* number of elements of the dataset */
int64_t nnx_h5_count(const char* path, const char* name, char* err, size_t errlen);
/* the whole dataset, row-major, converted to out_type as a C cast would;
 * returns the element count (at most cap) */
int64_t nnx_h5_read(const char* path, const char* name, int out_type, void* out, int64_t cap, char* err,
                    size_t errlen);

/* FileNetwork::GenerateFromHDF's sampling (src/network_factory.cu:171-212):
 * with sparsity = numpnts / N, neuron i is row i*sparsity of the file's CSR
 * (row_ptr: numpnts+1 entries) and keeps the targets divisible by sparsity,
 * divided by it.  Targets that land at N or above (numpnts not a multiple of
 * N; the reference keeps them, out of range) are dropped and counted in
 * *dropped.  offsets: N+1; targets: room for row_ptr[numpnts] entries.
 * Returns the number of edges, or -1 if N < 1 or N > numpnts (the
 * reference prints an error and divides by zero).                         */
int64_t nnx_csr_subsample(int64_t numpnts, const int32_t* row_ptr, const int32_t* col_idx, int64_t N,
                          int32_t* offsets, int32_t* targets, int64_t* dropped);

/* FileMorphology::GenerateFromHDF's placement (src_wholebrain/
 * morphology.cpp:301-370): neuron i at point (numpnts/N)*i of the axes,
 * mapped per axis to [-1, 1] by the bounding box of all numpnts points
 * (p - min, times 2/(max - min), minus 1, in float).  pos: 3N interleaved.
 * Returns 0, or -1 if N < 1 or N > numpnts.                                */
int nnx_morphology_from_axes(int64_t numpnts, const float* x, const float* y, const float* z, int64_t N,
                             float* pos);

/* Csr_matrix::input_from_connection_file (src/paracrine.cuh:49-108): the
 * text format "rows nnz" / rowPtr (rows+1) / colIdx (nnz) / val (nnz), one
 * line each.  With row_ptr NULL only *rows and *nnz are set.  Returns 0, or
 * -1 (message in err) for a missing file or a short line (the reference
 * prints a message and leaves the entries zero).                          */
int64_t nnx_read_connection_file(const char* path, int64_t* rows, int64_t* nnz, int32_t* row_ptr, int32_t* col,
                                 float* val, char* err, size_t errlen);

#ifdef __cplusplus
}
#endif
#endif /* NNX_HOST_H */
