// MT19937 pseudo-random generator (Matsumoto & Nishimura, ACM TOMACS 8(1), 1998),
// with the published init_by_array seeding (2002 revision) and the [0,1]
// closed-interval real draw. QuEST seeds and draws measurement outcomes with
// exactly this generator (QuEST/src/mt19937ar.c:62-155), so identical seeds
// give identical measurement outcomes. One global state, as in the reference.
#include "mt19937.h"

#include <cstdint>

namespace qmx {
namespace mt {

namespace {
constexpr int N = 624, M = 397;
constexpr uint32_t kMatrixA = 0x9908b0dfu, kUpper = 0x80000000u, kLower = 0x7fffffffu;
uint32_t g_state[N];
int g_index = N + 1;  // N+1: not seeded

void seed(uint32_t s) {
    g_state[0] = s;
    for (int i = 1; i < N; ++i)
        g_state[i] = 1812433253u * (g_state[i - 1] ^ (g_state[i - 1] >> 30)) + (uint32_t)i;
    g_index = N;
}

void twist() {
    for (int k = 0; k < N; ++k) {
        uint32_t y = (g_state[k] & kUpper) | (g_state[(k + 1) % N] & kLower);
        uint32_t v = g_state[(k + M) % N] ^ (y >> 1);
        g_state[k] = (y & 1u) ? (v ^ kMatrixA) : v;
    }
    g_index = 0;
}
}  // namespace

void initByArray(const unsigned long *key, int len) {
    seed(19650218u);
    int i = 1, j = 0;
    for (int k = (N > len ? N : len); k; --k) {
        g_state[i] = (g_state[i] ^ ((g_state[i - 1] ^ (g_state[i - 1] >> 30)) * 1664525u)) +
                     (uint32_t)key[j] + (uint32_t)j;
        ++i;
        ++j;
        if (i >= N) {
            g_state[0] = g_state[N - 1];
            i = 1;
        }
        if (j >= len) j = 0;
    }
    for (int k = N - 1; k; --k) {
        g_state[i] = (g_state[i] ^ ((g_state[i - 1] ^ (g_state[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
        ++i;
        if (i >= N) {
            g_state[0] = g_state[N - 1];
            i = 1;
        }
    }
    g_state[0] = 0x80000000u;
    g_index = N;
}

uint32_t nextU32() {
    if (g_index >= N) {
        if (g_index == N + 1) seed(5489u);
        twist();
    }
    uint32_t y = g_state[g_index++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

double nextReal1() { return nextU32() * (1.0 / 4294967295.0); }

}  // namespace mt
}  // namespace qmx

extern "C" {
// test hooks: expose the raw generator so its stream can be pinned against
// std::mt19937 and the reference's golden draws
void qmxMtInitByArray(const unsigned long *key, int len) { qmx::mt::initByArray(key, len); }
unsigned int qmxMtNextU32(void) { return qmx::mt::nextU32(); }
double qmxMtNextReal1(void) { return qmx::mt::nextReal1(); }
}
