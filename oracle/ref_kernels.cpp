/*
 * Instantiates the reference's own CPU kernel macros with exported C names,
 * so that oracle/_ref/liboskar_ref.so carries the reference's arithmetic for
 * the parts of the hot path whose CPU bodies live in macro headers.  This
 * file holds no arithmetic of its own: every body below is expanded from the
 * header named beside it, read from the reference tree at build time.
 */
#include "oskar_global.h"
#include "utility/oskar_vector_types.h"
#include "utility/oskar_kernel_macros.h"
#include "math/define_multiply.h"
#include "math/define_dftw_c2c.h"                     /* :79-109 */
#include "math/define_dftw_m2m.h"                     /* :109-167 */
#include "telescope/station/define_evaluate_element_weights_dft.h"
#include "telescope/station/element/define_evaluate_dipole_pattern.h"
#include "convert/define_convert_enu_directions_to_theta_phi.h"
#include "interferometer/define_evaluate_jones_K.h"   /* :56-75 */
#include "correlate/define_correlate_utils.h"
#include "correlate/define_auto_correlate.h"        /* :32-58, :84-104 */
#include "sky/oskar_sky_enum.h"
#include "sky/define_update_horizon_mask.h"          /* :3-12 */
#include "sky/define_sky_copy_source_data.h"         /* :3-25 */

#undef KERNEL
#define KERNEL(NAME) extern "C" OSKAR_EXPORT void NAME

OSKAR_DFTW_C2C_CPU(ref_dftw_c2c_2d_float, 0, float, float2)
OSKAR_DFTW_C2C_CPU(ref_dftw_c2c_3d_float, 1, float, float2)
OSKAR_DFTW_M2M_CPU(ref_dftw_m2m_2d_float, 0, float, float2)
OSKAR_DFTW_M2M_CPU(ref_dftw_m2m_3d_float, 1, float, float2)
OSKAR_DFTW_C2C_CPU(ref_dftw_c2c_2d_double, 0, double, double2)
OSKAR_DFTW_C2C_CPU(ref_dftw_c2c_3d_double, 1, double, double2)
OSKAR_DFTW_M2M_CPU(ref_dftw_m2m_2d_double, 0, double, double2)
OSKAR_DFTW_M2M_CPU(ref_dftw_m2m_3d_double, 1, double, double2)
OSKAR_ELEMENT_WEIGHTS_DFT(ref_element_weights_dft_double, double, double2)
OSKAR_ELEMENT_WEIGHTS_DFT(ref_element_weights_dft_float, float, float2)
OSKAR_EVALUATE_DIPOLE_PATTERN(ref_dipole_pattern_double, double, double2,
        double4c)
OSKAR_EVALUATE_DIPOLE_PATTERN_SCALAR(ref_dipole_pattern_scalar_double,
        double, double2)
OSKAR_CONVERT_ENU_DIR_TO_THETA_PHI(ref_theta_phi_double, double)
OSKAR_JONES_K_CPU(ref_jones_K_double, double, double2)
OSKAR_ACORR_CPU(ref_acorr_float, float, float2, float4c)
OSKAR_ACORR_CPU(ref_acorr_double, double, double2, double4c)
OSKAR_ACORR_SCALAR_CPU(ref_acorr_scalar_float, float, float2)
OSKAR_ACORR_SCALAR_CPU(ref_acorr_scalar_double, double, double2)
OSKAR_UPDATE_HORIZON_MASK(ref_update_horizon_mask_float, float)
OSKAR_UPDATE_HORIZON_MASK(ref_update_horizon_mask_double, double)
OSKAR_SKY_COPY_SOURCE_DATA(ref_sky_copy_source_data_float, float)
OSKAR_SKY_COPY_SOURCE_DATA(ref_sky_copy_source_data_double, double)
