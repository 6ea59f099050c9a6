// ssx_ax_f64.hip -- element operator instantiations, fields double,
// geometric factors double (template in ssx_ax.h).
#include "ssx_ax.h"

extern "C" int ssxk_ax_dispatch_f64(void *stream, int n, int v, int nel,
		int op, const void *u, void *au, const void *gf, const void *mass,
		const void *dmat, double lambda, const int *bc, int apply_mask,
		const int *done, double *part, const void *pu)
{
	if (nel <= 0)
		return SSX_OK;
	return ax_dispatch<double, double>((hipStream_t) stream, n, v, nel, op, u, au,
			gf, mass, dmat, lambda, bc, apply_mask, done, part,
			(const PUpd *) pu);
}
