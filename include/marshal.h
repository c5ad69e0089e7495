/*
 * marshal.h - the reference's CUDA-side entry points (include/marshal.h:26-31), kept with
 * identical names and signatures so existing callers relink without edits.
 *
 * `src` is a device pointer (hipMalloc'd), work goes to the null stream and is asynchronous
 * with respect to the host, as the reference's launches on the default stream are
 * (src/marshal.cu:27-80). `timer` is accepted and ignored, as the reference kernels ignore
 * it. Return value: false on success, true on error (diagnostic on stderr).
 */
#ifndef LIBMARSHAL_AMD_MARSHAL_H_
#define LIBMARSHAL_AMD_MARSHAL_H_

#include <time.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* AoS [height][width] -> ASTA [height/tile_size][width][tile_size], LDS-staged.
 * Replaces gpu_aos_asta_bs_float / _double, include/marshal.h:26-27 (src/marshal.cu:42-49). */
bool gpu_aos_asta_bs_float(float *src, int height, int width, int tile_size, clock_t *timer);
bool gpu_aos_asta_bs_double(double *src, int height, int width, int tile_size, clock_t *timer);

/* Same conversion by cycle following with LDS bit flags.
 * Replaces gpu_aos_asta_pttwac, include/marshal.h:29 (src/marshal.cu:52-62). */
bool gpu_aos_asta_pttwac(float *src, int height, int width, int tile_size, clock_t *timer);

/* SoA [width][height] -> ASTA [height/tile_size][width][tile_size] by cycle following with
 * global bit flags. Replaces gpu_soa_asta_pttwac, include/marshal.h:31 (src/marshal.cu:64-80). */
bool gpu_soa_asta_pttwac(float *src, int height, int width, int tile_size, clock_t *timer);

#ifdef __cplusplus
}
#endif

#endif /* LIBMARSHAL_AMD_MARSHAL_H_ */
