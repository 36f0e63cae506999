// Run-time compilation of a problem's user_source (hipRTC).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace amr {

// Compiles `code` (a __device__ user_source(i, j, k, state, rhs, pparm, cls,
// dx) definition) with the prelude and the per-cell kernel for `device`;
// 0 on success, else the compiler's log in *log.
int compile_user_source(const char* code, int device, hipModule_t* mod, hipFunction_t* fn, std::string* log);
// The definitions the user's code is compiled against.
const char* user_source_prelude();

}  // namespace amr
