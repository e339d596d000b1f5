/*
 * Device functions of the dipole element pattern, shared by the pattern
 * kernels (beam_util.hip) and the fused aperture-array beam (dftw.hip):
 * theta and phi of an ENU direction
 * (oskar/convert/define_convert_enu_directions_to_theta_phi.h) and the
 * pattern of a pair of crossed dipoles
 * (oskar/telescope/station/element/define_evaluate_dipole_pattern.h).
 */
#ifndef OSKAR_AMD_DIPOLE_DEVICE_H_
#define OSKAR_AMD_DIPOLE_DEVICE_H_

#include "device_common.h"

namespace oskar_amd {

/* theta and the two phi (each offset by its delta, in [0, 2 pi)) of one
 * ENU direction. */
template <typename FP>
__device__ __forceinline__ void theta_phi_point(FP xx, FP yy, FP zz,
        FP delta_phi1, FP delta_phi2, FP* theta, FP* phi1, FP* phi2)
{
    const FP twopi = 2 * pi_v<FP>();
    const FP p0 = atan2(yy, xx);
    FP p1 = fmod(p0 + delta_phi1, twopi);
    FP p2 = fmod(p0 + delta_phi2, twopi);
    if (p1 < (FP) 0) p1 += twopi;
    if (p2 < (FP) 0) p2 += twopi;
    *theta = atan2(sqrt(xx * xx + yy * yy), zz);
    *phi1 = p1;
    *phi2 = p2;
}

/* One dipole: E_theta, E_phi (imaginary parts zero). */
template <typename FP>
__device__ __forceinline__ void dipole(FP phi, FP sin_theta, FP cos_theta,
        FP kL, FP cos_kL, Cplx<FP>* e_theta, Cplx<FP>* e_phi)
{
    FP sin_phi, cos_phi;
    sincos_v<FP>(phi, &sin_phi, &cos_phi);
    const FP denom = (FP) 1 + cos_phi * cos_phi * (cos_theta * cos_theta -
            (FP) 1);
    e_theta->y = e_phi->y = (FP) 0;
    if (denom == (FP) 0)
    {
        e_theta->x = e_phi->x = (FP) 0;
    }
    else
    {
        const FP t = (cos(kL * cos_phi * sin_theta) - cos_kL) / denom;
        e_theta->x = -cos_phi * cos_theta * t;
        e_phi->x = sin_phi * t;
    }
}

/* The dipole pattern of one direction, written to pattern[o]: the 2x2
 * matrix (X theta, X phi; Y theta, Y phi), rows swapped with swap_xy, or
 * for a scalar pattern the amplitude sqrt(sum of squares / 2). */
template <typename FP, bool MATRIX>
__device__ __forceinline__ void dipole_point(FP theta, FP phi_x, FP phi_y,
        FP kL, FP cos_kL, int swap_xy, Cplx<FP>* pattern, size_t o)
{
    FP sin_theta, cos_theta;
    sincos_v<FP>(theta, &sin_theta, &cos_theta);
    Cplx<FP> x_th, x_ph, y_th, y_ph;
    dipole<FP>(phi_x, sin_theta, cos_theta, kL, cos_kL, &x_th, &x_ph);
    dipole<FP>(phi_y, sin_theta, cos_theta, kL, cos_kL, &y_th, &y_ph);
    if (MATRIX)
    {
        Cplx<FP>* p = pattern + 4 * o;
        if (swap_xy)
        {
            p[0] = y_th; p[1] = y_ph; p[2] = x_th; p[3] = x_ph;
        }
        else
        {
            p[0] = x_th; p[1] = x_ph; p[2] = y_th; p[3] = y_ph;
        }
    }
    else
    {
        const FP amp = sqrt((x_th.x * x_th.x + x_ph.x * x_ph.x +
                y_th.x * y_th.x + y_ph.x * y_ph.x) / (FP) 2);
        pattern[o].x = amp;
        pattern[o].y = (FP) 0;
    }
}

} /* namespace oskar_amd */

#endif /* OSKAR_AMD_DIPOLE_DEVICE_H_ */
