/* rgf_plain.h -- plain per-energy GPU form of green_RGF_RS (src/green_rgf.f90:3950-4164).
 *
 * The comparison baseline of the batched engine in rgf.h: energies are solved
 * one after another, each as a chain of rocBLAS zgemm and rocSOLVER
 * zgetrf/zgetri calls in the reference's own order, with the decimation's
 * convergence tested on the host after every iteration.  Same rgf_params and
 * rgf_batch contract (device arrays) as rgf_solve_batched; synchronous. */
#ifndef LITEQT_RGF_PLAIN_H
#define LITEQT_RGF_PLAIN_H
#include "rgf.h"
#ifdef __cplusplus
extern "C" {
#endif

int rgf_plain_solve(const rgf_params *p, const rgf_batch *b, void *stream);
const char *rgf_plain_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
