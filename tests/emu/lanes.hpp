// Test infrastructure: the lanes of one emulated workgroup as cooperative
// contexts on the calling thread.  Each lane runs the kernel on a stack of
// its own until it reaches a barrier (emu::barrier) or returns; the
// scheduler resumes the lanes in the phase's order (forward, backward or a
// seeded shuffle, so that a read of another lane's write that is not
// separated by a barrier gives order-dependent results) and checks that
// either every lane waits at a barrier or every lane has returned.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <random>
#include <vector>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define EMU_ASAN 1
#endif

#include "hip/hip_runtime.h"

// emu_switch(&save_sp, next_sp): pushes the callee-saved registers, stores
// the stack pointer in *save_sp, loads next_sp and pops that context's.
extern "C" void emu_switch(void **save_sp, void *next_sp);
asm(R"(
    .text
    .globl emu_switch
    .type emu_switch,@function
emu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size emu_switch, .-emu_switch
)");

namespace emu {

struct Lane {
    void *sp = nullptr;
    std::vector<unsigned char> stack;
    bool done = false;
};

struct Group {
    std::vector<Lane> lanes;
    void *sched_sp = nullptr;
    int cur = -1;
    std::function<void()> body;
    uint64_t phases = 0;
#ifdef EMU_ASAN
    void *fake = nullptr;
    const void *sched_bottom = nullptr;
    size_t sched_size = 0;
#endif
};

inline Group *g_group = nullptr;

inline void to_sched() {
    Group &g = *g_group;
#ifdef EMU_ASAN
    __sanitizer_start_switch_fiber(&g.fake, g.sched_bottom, g.sched_size);
#endif
    emu_switch(&g.lanes[g.cur].sp, g.sched_sp);
#ifdef EMU_ASAN
    __sanitizer_finish_switch_fiber(g.fake, nullptr, nullptr);
#endif
}

inline void barrier() { to_sched(); }

inline void lane_entry() {
#ifdef EMU_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &g_group->sched_bottom, &g_group->sched_size);
#endif
    g_group->body();
    g_group->lanes[g_group->cur].done = true;
    to_sched();
    std::abort();  // a finished lane is never resumed
}

// Runs body() on `lanes` lanes (threadIdx.x = lane) to completion.  order:
// 0 forward, 1 backward, >= 2 a shuffle seeded by `order` for every phase.
inline void run_group(int lanes, int order, std::function<void()> body, size_t stack_bytes = 256 << 10) {
    Group g;
    g.body = std::move(body);
    g.lanes.resize(lanes);
    for (auto &L : g.lanes) {
        L.stack.resize(stack_bytes);
        uintptr_t top = ((uintptr_t)(L.stack.data() + L.stack.size())) & ~(uintptr_t)15;
        void **sp = (void **)top;
        *--sp = nullptr;                        // lane_entry's return slot (never used)
        *--sp = (void *)&lane_entry;            // where emu_switch returns to
        for (int i = 0; i < 6; ++i) *--sp = nullptr;  // rbp rbx r12 r13 r14 r15
        L.sp = sp;
    }
    g_group = &g;
    std::mt19937_64 rng((uint64_t)order);
    std::vector<int> idx(lanes);
    for (int i = 0; i < lanes; ++i) idx[i] = i;
    while (true) {
        if (order == 1) std::reverse(idx.begin(), idx.end());
        else if (order >= 2) std::shuffle(idx.begin(), idx.end(), rng);
        for (int i : idx) {
            if (g.lanes[i].done) continue;
            g.cur = i;
            thread_idx.x = (unsigned)i;
#ifdef EMU_ASAN
            const Lane &L = g.lanes[i];
            __sanitizer_start_switch_fiber(&g.fake, L.stack.data(), L.stack.size());
#endif
            emu_switch(&g.sched_sp, g.lanes[i].sp);
#ifdef EMU_ASAN
            __sanitizer_finish_switch_fiber(g.fake, nullptr, nullptr);
#endif
        }
        if (order == 1) std::reverse(idx.begin(), idx.end());
        int done = 0;
        for (const auto &L : g.lanes) done += L.done;
        ++g.phases;
        if (done == lanes) break;
        if (done != 0) {
            std::fprintf(stderr, "lanes diverged at a barrier: %d of %d returned in phase %llu\n", done, lanes,
                         (unsigned long long)g.phases);
            std::abort();
        }
    }
    g_group = nullptr;
}

}  // namespace emu
