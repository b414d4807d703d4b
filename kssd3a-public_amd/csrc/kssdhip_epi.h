// FP64 value of one (query, reference) cell from its counts, as the matrix
// outputs compute it (make_selected_output_row through
// fill_row_base_distances, command_ani.c:1321-1379, for the metrics that
// need no calibration model).  Shared by the dense epilogue and the
// query-major select kernel.
#ifndef KSSDHIP_EPI_H
#define KSSDHIP_EPI_H
#include <hip/hip_runtime.h>

#include "kssd_models.h"
#include "kssdhip.h"

namespace kssd {

// selected_metric_effective_id without calibration (command_ani.c:1260-1270)
__device__ __forceinline__ int epi_effective_metric(const kssdhip_epilogue_opts_t &o)
{
	int s = o.metric;
	if (s < 1 || s > 9)
		s = 1;
	if (o.unassembled && s <= 4)
		return 4;
	return s;
}

// qc / rc: the query's and the reference's context totals (both > 0)
__device__ __forceinline__ double epi_cell_value(const kssdhip_epilogue_opts_t &o, int s, uint32_t xny, uint32_t mut2,
						 uint32_t dobj, uint32_t dsec, uint32_t qc, uint32_t rc)
{
	const uint32_t k = o.ctx_bits / 2;
	double afq = (double)xny / (double)qc, afr = (double)xny / (double)rc;
	if (o.float_af) {
		afq = (double)((float)xny / (float)qc);
		afr = (double)((float)xny / (float)rc);
	}
	const bool af_pass = (afq > afr ? afq : afr) >= o.afcut;
	double dist;
	switch (s) {
	case 4:
		dist = af_pass ? kssd_naive_dist(xny, dobj, dsec) : o.exception;
		break;
	case 5:
		dist = kssd_mashD(k, rc, qc, xny);
		break;
	case 6:
		dist = kssd_aafD(k, rc, qc, xny);
		break;
	case 7:
		dist = af_pass ? kssd_moe_dist(xny, dobj, dsec, mut2, o.drfold) : kssd_mashD(k, rc, qc, xny);
		break;
	case 8:
		dist = af_pass ? kssd_moe_dist(xny, dobj, dsec, mut2, o.drfold) : kssd_aafD(k, rc, qc, xny);
		break;
	case 9:
		dist = af_pass ? kssd_p_dist(xny, dsec, o.obj_bits) : o.exception;
		break;
	default: // 3 (and 1/2 where no calibration is available: CtxMoE)
		dist = af_pass ? kssd_moe_dist(xny, dobj, dsec, mut2, o.drfold) : o.exception;
		break;
	}
	return o.values_ani ? 1.0 - dist : dist;
}

// Whether the epilogue can evaluate this metric without the host's
// calibration model.
static inline bool epi_metric_supported(const kssdhip_epilogue_opts_t &o)
{
	if (o.metric >= 3 && o.metric <= 9)
		return true;
	return o.unassembled && o.metric >= 1 && o.metric <= 4;
}

} // namespace kssd
#endif
