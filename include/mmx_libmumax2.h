/* mmx_libmumax2.h -- libmumax2's C functions on the effective-field / LLG /
 * RK12 path with their original argument lists, exported by libmmx.so under
 * the mmx_ prefix.
 *
 * Conventions (src/mumax/gpu/libmumax2.go and the src/libmumax2 headers):
 *  - every array is a float** holding one device pointer per used GPU; this
 *    library uses one GPU (mmx_setUsedGPUs with Ngpu == 1), so element [0];
 *  - a vector field is passed as its 3 component arrays (SoA, internal
 *    (z,y,x) order); a map argument whose float** or element [0] is NULL is
 *    uniform (1) and its scalar multiplier carries the value;
 *  - sizes are per GPU (Npart = floats per component), streams are an array
 *    of hipStream_t (the reference's CUstream*), one per used GPU;
 *  - the calls are asynchronous on that stream and return nothing: an
 *    invalid argument or a HIP error prints a message and aborts the
 *    process, as libmumax2/gpu_safe.h:36-41 does.
 * A cgo wrapper of mumax/gpu keeps its argument marshalling and changes only
 * the C symbol (C.torqueAsync -> C.mmx_torqueAsync). */
#ifndef MMX_LIBMUMAX2_H
#define MMX_LIBMUMAX2_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/libmumax2/multigpu.h:17 setUsedGPUs (one GPU: Ngpu must be 1) */
void mmx_setUsedGPUs(int *gpuIds, int Ngpu);

/* src/libmumax2/normalize.h:22 normalizeAsync */
void mmx_normalizeAsync(float **mx, float **my, float **mz, float **norm_map,
                        hipStream_t *stream, int Npart);

/* src/libmumax2/torque.h:35 torqueAsync */
void mmx_torqueAsync(float **tx, float **ty, float **tz, float **mx, float **my,
                     float **mz, float **hx, float **hy, float **hz,
                     float **alpha_map, float alpha_mul, hipStream_t *stream,
                     int Npart);

/* src/libmumax2/exchange6.h:20 exchange6Async */
void mmx_exchange6Async(float **hx, float **hy, float **hz, float **mx,
                        float **my, float **mz, float **msat, float **aex,
                        float Aex2_mu0MsatMul, int N0, int N1Part, int N2,
                        int periodic0, int periodic1, int periodic2,
                        float cellSizeX, float cellSizeY, float cellSizeZ,
                        hipStream_t *streams);

/* src/libmumax2/uniaxialanisotropy.h:20-26 uniaxialAnisotropyAsync */
void mmx_uniaxialAnisotropyAsync(float **hx, float **hy, float **hz,
                                 float **mx, float **my, float **mz,
                                 float **Ku1_map, float **MSat_map,
                                 float Ku2_Mu0Msat_mul, float **anisU_mapx,
                                 float anisU_mulx, float **anisU_mapy,
                                 float anisU_muly, float **anisU_mapz,
                                 float anisU_mulz, hipStream_t *stream,
                                 int Npart);

/* src/libmumax2/zhang-li_torque.h:21-35 zhangli_async */
void mmx_zhangli_async(float **sttx, float **stty, float **sttz, float **mx,
                       float **my, float **mz, float **jx, float **jy,
                       float **jz, float **msat, float **pol, float **ee,
                       float **alpha, const float jMul_x, const float jMul_y,
                       const float jMul_z, const float pred, const float eeMul,
                       const float alphaMul, const int sx, const int sy,
                       const int sz, const float csx, const float csy,
                       const float csz, const int pbc_x, const int pbc_y,
                       const int pbc_z, hipStream_t *stream);

/* src/libmumax2/add.h:45 maddAsync: dst = a + mulB*b (b NULL: a + mulB) */
void mmx_maddAsync(float **dst, float **a, float **b, float mulB,
                   hipStream_t *stream, int Npart);
/* src/libmumax2/add.h:36 madd1Async: a += mulB*b */
void mmx_madd1Async(float **a, float **b, float mulB, hipStream_t *stream,
                    int Npart);
/* src/libmumax2/add.h:40 madd2Async: a += mulB*b + mulC*c */
void mmx_madd2Async(float **a, float **b, float mulB, float **c, float mulC,
                    hipStream_t *stream, int Npart);
/* src/libmumax2/add.h:49-54 vecMaddAsync: dst_j = a_j + mulB_j*b_j */
void mmx_vecMaddAsync(float **dstx, float **dsty, float **dstz, float **ax,
                      float **ay, float **az, float **bx, float **by,
                      float **bz, float mulBx, float mulBy, float mulBz,
                      hipStream_t *stream, int Npart);
/* src/libmumax2/mul.h:20 mulAsync: dst = a*b */
void mmx_mulAsync(float **dst, float **a, float **b, hipStream_t *stream,
                  int Npart);
/* src/libmumax2/add.h:61 cmaddAsync: dst += (a + b i) kern src
 * (dst, src interleaved complex; kern real; NComplexPart complex values) */
void mmx_cmaddAsync(float **dst, float a, float b, float **kern, float **src,
                    hipStream_t *stream, int NComplexPart);

/* src/libmumax2/kernelmul_micromag.h:25-28 kernelMulMicromag3DAsync: the
 * in-place 3x3 symmetric real-kernel product of ConvPlan.Convolve
 * (gpu/convplan.go:163-176); partLen3D floats per fftM component, half as
 * many per kernel component */
void mmx_kernelMulMicromag3DAsync(float **fftMx, float **fftMy, float **fftMz,
                                  float **fftKxx, float **fftKyy,
                                  float **fftKzz, float **fftKyz,
                                  float **fftKxz, float **fftKxy,
                                  hipStream_t *stream, int partLen3D);

/* src/libmumax2/copypad.h:37 copyPad3DAsync / :39 copyUnPad3DAsync: copy the
 * (S0,S1,S2) block into the corner of the (D0,D1,D2) array (pad; the rest of
 * dst is not written) or the (D0,D1,D2) corner of src out (unpad), Ncomp
 * components stored one after another */
void mmx_copyPad3DAsync(float **dst, int D0, int D1, int D2, float **src,
                        int S0, int S1, int S2, int Ncomp,
                        hipStream_t *streams);
void mmx_copyUnPad3DAsync(float **dst, int D0, int D1, int D2, float **src,
                          int S0, int S1, int S2, int Ncomp,
                          hipStream_t *streams);
/* src/libmumax2/copypad.h:49 zeroArrayAsync */
void mmx_zeroArrayAsync(float **A, int N, hipStream_t *streams);

/* src/libmumax2/reduce.h:34-119: one float partial per block in
 * output[0][0 .. blocks-1], finished on the host by the caller
 * (gpu/reductor.go:19-212).  Sums accumulate in FP64 within a block and are
 * rounded once; extrema are exact.  threadsPerBlock is accepted and not
 * used (the workgroup size is this library's own). */
void mmx_partialSumAsync(float **input, float **output, int blocksPerGPU,
                         int threadsPerBlockPerGPU, int NPerGPU,
                         hipStream_t *streams);
void mmx_partialSDotAsync(float **input1, float **input2, float **output,
                          int blocks, int threadsPerBlock, int N,
                          hipStream_t *stream);
void mmx_partialMaxAsync(float **input, float **output, int blocksPerGPU,
                         int threadsPerBlockPerGPU, int NPerGPU,
                         hipStream_t *streams);
void mmx_partialMinAsync(float **input, float **output, int blocksPerGPU,
                         int threadsPerBlockPerGPU, int NPerGPU,
                         hipStream_t *streams);
void mmx_partialMaxAbsAsync(float **input, float **output, int blocksPerGPU,
                            int threadsPerBlockPerGPU, int NPerGPU,
                            hipStream_t *streams);
void mmx_partialMaxDiffAsync(float **a, float **b, float **output,
                             int blocksPerGPU, int threadsPerBlockPerGPU,
                             int NPerGPU, hipStream_t *streams);
void mmx_partialMaxSumAsync(float **a, float **b, float **output,
                            int blocksPerGPU, int threadsPerBlockPerGPU,
                            int NPerGPU, hipStream_t *streams);
void mmx_partialMaxNorm3SqAsync(float **x, float **y, float **z,
                                float **output, int blocksPerGPU,
                                int threadsPerBlockPerGPU, int NPerGPU,
                                hipStream_t *streams);
void mmx_partialMaxNorm3SqDiffAsync(float **x1, float **y1, float **z1,
                                    float **x2, float **y2, float **z2,
                                    float **output, int blocksPerGPU,
                                    int threadsPerBlockPerGPU, int NPerGPU,
                                    hipStream_t *streams);

#ifdef __cplusplus
}
#endif

#endif /* MMX_LIBMUMAX2_H */
