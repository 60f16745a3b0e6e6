// opo_internal.hpp -- shared host helpers of the library (not part of the ABI).
#pragma once
#include <cstdarg>
#include <cstddef>
#include <cstdint>

#include "../../include/opo3d.h"

namespace opo {

// Record a formatted error for opo_last_error() and return code.
int fail(int code, const char *fmt, ...);
void clear_error();

// omega_j of the reference rule (EFields3D.cuh:318-328; Common.cuh:120-128;
// EFields3D.cuh:58-68): 2 pi fftshift(linspace(-NT/(2W), NT/(2W), NT)).
void omega_rule(int nt, double W, double *w);

// exp(i phi) with phi reduced mod 2 pi in FP64 before cos/sin.
void cis(double phi, double &c, double &s);

}  // namespace opo
