/*
 * hb_plan.h -- what other translation units of the library may read of a
 * plan (hb_host.c owns the struct): its device record, stream and device.
 */
#ifndef HB_PLAN_H
#define HB_PLAN_H

#include <hip/hip_runtime_api.h>
#include "hb_shared.h"
#include "../../include/hbicg.h"

#ifdef __cplusplus
extern "C" {
#endif

const hb_dev *hb_plan_dev(const hbicg_plan *pl);
hipStream_t   hb_plan_stream(const hbicg_plan *pl);
int           hb_plan_device(const hbicg_plan *pl);

#ifdef __cplusplus
}
#endif
#endif
