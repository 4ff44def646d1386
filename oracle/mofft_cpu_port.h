/*
 * mofft_cpu_port.h - the reference executor on one host core, with every
 * data-independent table made at plan time (timed as bench.py's CPU
 * baseline).  TEST AND MEASUREMENT INFRASTRUCTURE ONLY: the product library
 * never links or calls it.
 *
 * Arrays are interleaved complex ({re, im} pairs), one transform per call;
 * `sign` is -1 forward, +1 backward; `radices` run outermost first
 * (include/mofft.h:138-139).  create returns NULL for a chain the reference
 * rejects (src/mofft.c:27-28, 353-417).  A plan's scratch is its own: one
 * plan must not execute on two threads at once (as in the reference,
 * src/mofft.c:1944-1945).
 */
#ifndef MOFFT_CPU_PORT_H
#define MOFFT_CPU_PORT_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct port_plan_f32 port_plan_f32;
typedef struct port_plan_f64 port_plan_f64;

port_plan_f32 *port_create_f32(size_t n, int sign, const int *radices,
                               size_t depth);
port_plan_f64 *port_create_f64(size_t n, int sign, const int *radices,
                               size_t depth);
int port_execute_f32(const port_plan_f32 *plan, const float *in, float *out);
int port_execute_f64(const port_plan_f64 *plan, const double *in, double *out);
void port_destroy_f32(port_plan_f32 *plan);
void port_destroy_f64(port_plan_f64 *plan);

#ifdef __cplusplus
}
#endif
#endif
