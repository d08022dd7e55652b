// ORACLE — test infrastructure only: shared includes of the CPU restatements under oracle/. Only
// tests/, __graft_entry__.smoke() and bench.py's CPU baseline load the library built from here.
#pragma once
#include <cstdint>

#include "caelix_amd/layout.h"
