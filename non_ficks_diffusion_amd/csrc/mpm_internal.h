// Internal declarations shared by the host-side sources of libmpm_gpu.
#pragma once
#include <cstdint>
#include <string>

int mpm_set_error(int code, const std::string& msg);
