/*
 * ORACLE (test infrastructure only): FP64 path-tracing reference images.
 * Filled in with the wavefront path tracer's estimator.
 */
int or_pt64_available(void) { return 0; }
