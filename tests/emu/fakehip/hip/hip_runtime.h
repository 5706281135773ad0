// Test infrastructure only: the emulation build's stand-in for <hip/hip_runtime.h>.
#pragma once
#include "../../hipemu.hpp"
