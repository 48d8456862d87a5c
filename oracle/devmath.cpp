/*
 * CPU ORACLE — test infrastructure only (see mc_oracle.h).
 *
 * The device's fp32 power and logarithm (ddrs_amd/csrc/mc_math.h, which is
 * host and device code: every multiply-add is an explicit fma, so the host
 * build gives the device's bits) as C functions for the oracle's third
 * instantiation, orc_*_c: float arithmetic with these powers in place of
 * glibc's powf / logf. The reference's f32 path calls libm, which rounds a
 * small fraction of powers differently from this correctly rounded one; the
 * _c variant therefore separates the kernels' arithmetic (op order, CSR
 * order, clamps) from libm's rounding when the two oracles are compared with
 * the device on workloads where the recurrence amplifies single roundings.
 */
#include "../ddrs_amd/csrc/mc_math.h"

extern "C" float orc_devpowf(float x, float y) { return ddrs::pow_f32(x, y); }
extern "C" float orc_devlogf(float x) { return ddrs::log_f32(x); }
