// hf_noalloc.h -- the device scratch behind hf_step_noalloc (internal).
#pragma once
#include <string>

// Size of the per-GPU scratch array: enough for the tables and working arrays
// of, e.g., one 8192^2 block or 512 blocks of up to 512^2 cells.
#define HF_NOALLOC_SCRATCH_BYTES (32u << 20)
#define HF_NOALLOC_MAX_DEVICES 64

// Device address of the current device's scratch array (looked up with
// hipGetSymbolAddress on the first call per device, then cached).  Returns 0,
// or -1 with *err set.
int hf_noalloc_scratch(void** out, std::string* err);
