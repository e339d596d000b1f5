/*
 * Weighted DFT used by the aperture-array station beam.
 *
 * Reference: oskar/math/oskar_dftw.h:92-111, oskar/math/src/oskar_dftw.c:42-374
 * and the kernel bodies oskar/math/define_dftw_c2c.h, define_dftw_m2m.h:
 *
 *   out[offset_out + j] = norm * sum_i data[idx_i * num_out + j] * w_i
 *        * exp(+i k (x_o[j] x_i + y_o[j] y_i [+ z_o[j] z_i]))
 *
 * with x_o[j] = x_out[offset_coord_out + j], idx_i = data_idx[i] (or i when
 * data_idx is NULL), norm = 1/num_in if normalise.  The transform is 3D only
 * when both z_in and z_out are given.  For complex-matrix data, eval_x
 * selects the output .a/.b pair and eval_y the .c/.d pair; the other pair is
 * left untouched.  output is grown with oskar_mem_ensure to
 * offset_out + num_out.  OSKAR_GPU runs the HIP kernel on the current HIP
 * stream; OSKAR_CPU returns OSKAR_ERR_BAD_LOCATION.
 */
#ifndef OSKAR_DFTW_H_
#define OSKAR_DFTW_H_

#include "oskar_mem.h"

#ifdef __cplusplus
extern "C" {
#endif

OSKAR_EXPORT
void oskar_dftw(
        int normalise,
        int num_in,
        double wavenumber,
        const oskar_Mem* weights_in,
        const oskar_Mem* x_in,
        const oskar_Mem* y_in,
        const oskar_Mem* z_in,
        int offset_coord_out,
        int num_out,
        const oskar_Mem* x_out,
        const oskar_Mem* y_out,
        const oskar_Mem* z_out,
        const oskar_Mem* data_idx,
        const oskar_Mem* data,
        int eval_x,
        int eval_y,
        int offset_out,
        oskar_Mem* output,
        int* status);

#ifdef __cplusplus
}
#endif

#endif /* OSKAR_DFTW_H_ */
