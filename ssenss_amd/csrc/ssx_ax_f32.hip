// ssx_ax_f32.hip -- element operator instantiations, fields float,
// geometric factors float (template in ssx_ax.h).
#include "ssx_ax.h"

extern "C" int ssxk_ax_dispatch_f32(void *stream, int n, int v, int nel,
		int op, const void *u, void *au, const void *gf, const void *mass,
		const void *dmat, double lambda, const int *bc, int apply_mask,
		const int *done, double *part, const void *pu)
{
	if (nel <= 0)
		return SSX_OK;
	return ax_dispatch<float, float>((hipStream_t) stream, n, v, nel, op, u, au,
			gf, mass, dmat, lambda, bc, apply_mask, done, part,
			(const PUpd *) pu);
}
