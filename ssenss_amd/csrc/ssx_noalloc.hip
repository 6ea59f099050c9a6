// ssx_noalloc.hip -- the host-array particle step for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).
//
// ssx_particles_step_host stages the caller's host particle arrays through
// device memory the context grows on demand (hipMalloc when a larger list
// arrives).  This variant stages them through one file-scope device array
// instead, of a fixed size, and keeps its two counters (largest pass count,
// particles naming no element) there too.  Each GPU holds its own copy of
// the array from the first use of the library on it until the process ends.
// Every call on one GPU shares that copy: calls that overlap on one GPU
// race.  The counters are cleared by a fill kernel at the start of every
// call; the staging area is fully written by the uploads before the walk
// reads it.
#include <pthread.h>

#include <hip/hip_runtime.h>

#include "ssx_internal.h"

#define SSX_NOALLOC_HEAD 256    /* counters, padded to the array alignment */
#define SSX_NOALLOC_BYTES                                                    \
	(SSX_NOALLOC_HEAD + (size_t) SSX_NOALLOC_MAX_PARTICLES                   \
	 * (8 * sizeof(double) + 3 * sizeof(int)))
#define SSX_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[SSX_NOALLOC_BYTES];

// per-device address of `scratch` (the host-side name is only a stand-in)
static void *scratch_dev[SSX_NOALLOC_MAX_DEVICES];
static pthread_mutex_t scratch_lock = PTHREAD_MUTEX_INITIALIZER;

// address of this device's copy; device `dev` must be current
static int scratch_ptr(int dev, void **out)
{
	if (dev < 0 || dev >= SSX_NOALLOC_MAX_DEVICES) {
		ssx_set_error("no-allocation particle step: device %d beyond the "
				"%d supported", dev, SSX_NOALLOC_MAX_DEVICES);
		return SSX_EINVAL;
	}
	pthread_mutex_lock(&scratch_lock);
	int rc = SSX_OK;
	if (!scratch_dev[dev]) {
		void *p = NULL;
		hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p) {
			ssx_set_error("hipGetSymbolAddress(scratch): %s",
					hipGetErrorString(e));
			rc = SSX_EHIP;
		} else {
			scratch_dev[dev] = p;
		}
	}
	*out = scratch_dev[dev];
	pthread_mutex_unlock(&scratch_lock);
	return rc;
}

extern "C" int ssx_particles_step_host_noalloc(ssx_ctx *c,
		const ssx_particles *h, double t_now, double dt, int *max_passes)
{
	if (!c || !h) {
		ssx_set_error("ssx_particles_step_host_noalloc: null argument");
		return SSX_EINVAL;
	}
	if (h->n > SSX_NOALLOC_MAX_PARTICLES) {
		ssx_set_error("ssx_particles_step_host_noalloc: %ld particles, more "
				"than the %d this entry point stages (use "
				"ssx_particles_step_host)", h->n, SSX_NOALLOC_MAX_PARTICLES);
		return SSX_EINVAL;
	}
	int prev = -1;
	if (hipGetDevice(&prev) != hipSuccess) {
		ssx_set_error("hipGetDevice failed");
		return SSX_EHIP;
	}
	if (prev != c->device && hipSetDevice(c->device) != hipSuccess) {
		ssx_set_error("hipSetDevice(%d) failed", c->device);
		return SSX_EHIP;
	}
	void *base = NULL;
	int rc = scratch_ptr(c->device, &base);
	if (rc == SSX_OK) {
		unsigned char *b = (unsigned char *) base;
		rc = ssx_walk_host(c, h, t_now, dt, max_passes, 0,
				b + SSX_NOALLOC_HEAD, SSX_NOALLOC_BYTES - SSX_NOALLOC_HEAD,
				(int *) b);
	}
	if (prev != c->device && hipSetDevice(prev) != hipSuccess
			&& rc == SSX_OK) {
		ssx_set_error("hipSetDevice(%d) failed", prev);
		rc = SSX_EHIP;
	}
	return rc;
}
