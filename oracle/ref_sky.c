/*
 * Instantiates the reference's two 2D direction-conversion macros
 * (oskar/convert/define_convert_relative_directions_to_lon_lat.h and
 * define_convert_lon_lat_to_relative_directions.h) under the names that the
 * reference's own conversion files give them
 * (oskar/convert/src/oskar_convert_relative_directions_to_lon_lat.c:17-20,
 * oskar_convert_lon_lat_to_relative_directions.c:15-18), plus the 3D forms
 * under ref_ names.  Those files also hold the GPU dispatchers, which need
 * the reference's device layer; compiling the macros here gives
 * oracle/_ref/liboskar_sky_ref.so the reference's CPU arithmetic without
 * it.  This file holds no arithmetic of its own: it is compiled as C, as
 * the reference compiles those files, so the macro bodies keep C's float
 * to double promotions.
 */
#include "oskar_global.h"
#include "utility/oskar_kernel_macros.h"
#include "convert/define_convert_relative_directions_to_lon_lat.h"
#include "convert/define_convert_lon_lat_to_relative_directions.h"

#include <math.h>

OSKAR_CONVERT_REL_DIR_TO_LON_LAT(oskar_convert_relative_directions_to_lon_lat_2d_f, 0, float)
OSKAR_CONVERT_REL_DIR_TO_LON_LAT(oskar_convert_relative_directions_to_lon_lat_2d_d, 0, double)
OSKAR_CONVERT_LON_LAT_TO_REL_DIR(oskar_convert_lon_lat_to_relative_directions_2d_f, 0, float)
OSKAR_CONVERT_LON_LAT_TO_REL_DIR(oskar_convert_lon_lat_to_relative_directions_2d_d, 0, double)
OSKAR_CONVERT_LON_LAT_TO_REL_DIR(ref_lon_lat_to_relative_directions_3d_float, 1, float)
OSKAR_CONVERT_LON_LAT_TO_REL_DIR(ref_lon_lat_to_relative_directions_3d_double, 1, double)
