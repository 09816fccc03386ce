// Index types of the CSR graph (reference include/types.h:3-4).
#pragma once
#include <cstdint>
typedef int32_t v_index_t;
typedef int64_t e_index_t;
