/*
 * multi_bspline.h -- umbrella header (replaces src/multi_bspline.h).
 *
 * Declares the 3-D uniform multi-spline API: types, create/set, single-point
 * evaluation for s/d/c/z, and the device-table / batched API.
 */
#ifndef EINSPLINE_MULTI_BSPLINE_H
#define EINSPLINE_MULTI_BSPLINE_H

#include "bspline_base.h"
#include "multi_bspline_structs.h"
#include "multi_bspline_eval_s.h"
#include "multi_bspline_eval_c.h"
#include "multi_bspline_eval_d.h"
#include "multi_bspline_eval_z.h"
#include "multi_bspline_create.h"
#include "multi_bspline_hip.h"

#endif
