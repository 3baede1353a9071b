// Host side of the tabulated EoS (VH_EOS_TABLE): validation, natural cubic
// spline coefficients and the device copy used by DevEos.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/vhlle_gpu.h"
#include "vh_device.h"

namespace vh {

// The SplineTab layout of a table EoS (x, then y, b, c, d of p, T, mu); n rows.
// VH_ERR_ARG if the table is invalid (fewer than 3 rows, e not strictly
// increasing, null columns).
int eos_table_host(const vh_eos_t &eos, std::vector<double> &buf, int &n);

// DevEos of a descriptor; for a table, *dbuf receives a device buffer that
// the caller frees (hipFree) after the last kernel that uses it.  st: the
// status word the table evaluations report to (may be null).
int make_dev_eos(const vh_eos_t *eos, int *st, DevEos &de, double **dbuf);

}  // namespace vh
