/*
 * Wavefront path tracing frames (REND_MODE_PATHTRACE).
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>

#include "rend_int.h"
#include "rend_hip.h"
#include "librend.h"

int rl_render_pt(struct rl_scene *s, unsigned rw, unsigned rh, unsigned *rays)
{
	(void)s;
	(void)rw;
	(void)rh;
	*rays = 0;
	return rl_fail(REND_EINVAL, "path tracing mode is not available yet");
}
