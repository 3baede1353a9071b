// Host side of the tabulated EoSs (VH_EOS_TABLE, VH_EOS_CHIRAL, VH_EOS_HADRON):
// validation, natural cubic spline coefficients and the device copies used by
// DevEos.
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

// The scalars of a VH_EOS_CHIRAL or VH_EOS_HADRON EoS into de, its table
// pointers at the descriptor's host arrays; v, stat (may be null): the tables
// concatenated for a device copy (the double fields, the hadron node flags).
// VH_ERR_ARG for an invalid descriptor.
int charged_tables_host(const vh_eos_t &eos, DevEos &de, std::vector<double> *v,
                        std::vector<int> *stat);
// points de's tables into arrays laid out as charged_tables_host returns them
void point_tables(DevEos &de, const double *v, const int *stat);

}  // namespace vh
