// gold.hip -- per-device initialisation of the Gold jump-ahead tables.
#include <mutex>

#include "common.h"
#include "gold.h"

namespace airan {

__constant__ GoldJumpTable kGoldJump;

namespace {

uint32_t StepColumn(int lfsr, uint32_t state)
{
    uint32_t nb = lfsr == 0 ? ((state >> 3) ^ state) & 1u
                            : ((state >> 3) ^ (state >> 2) ^ (state >> 1) ^ state) & 1u;
    return (state >> 1) | (nb << 30);
}

uint32_t HostMatVec(const uint32_t *col, uint32_t v)
{
    uint32_t r = 0;
    for (int j = 0; j < 31; ++j)
        if ((v >> j) & 1u) r ^= col[j];
    return r;
}

std::mutex g_mutex;
bool g_done[64] = {};

}  // namespace

int InitGoldTables()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return AIRAN_LAUNCH_FAILED;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_done[dev]) return AIRAN_OK;
    GoldJumpTable t{};
    for (int s = 0; s < 2; ++s) {
        for (int j = 0; j < 31; ++j) t.col[s][0][j] = StepColumn(s, 1u << j);
        for (int k = 1; k < kGoldJumpBits; ++k)
            for (int j = 0; j < 31; ++j) t.col[s][k][j] = HostMatVec(t.col[s][k - 1], t.col[s][k - 1][j]);
    }
    if (hipMemcpyToSymbol(HIP_SYMBOL(kGoldJump), &t, sizeof(t)) != hipSuccess) return AIRAN_LAUNCH_FAILED;
    g_done[dev] = true;
    return AIRAN_OK;
}

}  // namespace airan

extern "C" int32_t airan_InitDevice(void)
{
    int32_t st = airan::InitGoldTables();
    if (st == AIRAN_OK) st = airan::InitOfdmTables();
    return st != AIRAN_OK ? st : airan::InitLdpcScratch();
}
