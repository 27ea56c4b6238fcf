// MT19937 used for measurement outcomes (see mt19937.cpp).
#pragma once

#include <cstdint>

namespace qmx {
namespace mt {
void initByArray(const unsigned long *key, int len);
uint32_t nextU32();
double nextReal1();  // [0, 1]
}  // namespace mt
}  // namespace qmx
