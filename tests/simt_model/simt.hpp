// simt.hpp -- a CPU model of one HIP workgroup, for checking kernels on the host (test infrastructure only).
//
// Every GPU thread of a workgroup runs as a fiber on one OS thread (x86-64 stack switch below: callee-saved
// registers pushed on the old stack, stack pointer exchanged).  A fiber runs until it reaches a
// synchronisation point and then yields:
//   * a workgroup barrier (__syncthreads): released when every thread of the workgroup waits at the SAME
//     barrier (same source file and line); a thread that waits at another barrier, at a wavefront operation,
//     or that has left the kernel while others wait, is reported as a barrier mismatch;
//   * a wavefront operation (DPP move, readlane, shuffle): released when all 64 lanes of the wavefront wait at
//     the same operation; the result of every lane is computed from the values all lanes deposited.
// When no fiber can run and the workgroup has not finished, the model reports a deadlock with the position of
// every thread.  Workgroups run one after another, so __shared__ variables (function-local statics) are the
// LDS of the workgroup that runs.  Built with AddressSanitizer, every out-of-bounds access to LDS (static
// arrays), to global buffers (exact-size heap blocks) or to scratch is reported with its source line.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

// simt_switch(save, to): saves the callee-saved registers on the current stack, stores the stack pointer in
// *save, switches to stack `to` and restores the registers saved there.  A weak symbol: every translation unit
// that includes this header emits the same definition and the linker keeps one.
extern "C" void simt_switch(void **save, void *to);
asm(R"(
    .text
    .weak simt_switch
    .type simt_switch, @function
simt_switch:
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
    .size simt_switch, .-simt_switch
)");

namespace simt
{
    constexpr int kWave = 64;

    enum class Wait
    {
        running,
        barrier,
        wave,
        done
    };

    // Kinds of wavefront operation.  Every lane deposits a 32-bit value; the model computes each lane's result.
    enum class WaveOp
    {
        dpp,
        readlane,
        shfl,
        ballot,   // 64-bit mask of the lanes that deposited a non-zero value
        mfma_f32, // 16x16x4 FP32 matrix multiply-accumulate (operands and accumulators in vin, result in vout)
        mfma_f64, // 16x16x4 FP64
    };

    struct Thread
    {
        void *sp = nullptr;
        std::unique_ptr<unsigned char[]> stack;
        int tid = 0;
        Wait state = Wait::running;
        const char *file = nullptr;
        int line = 0;
        // wavefront operation
        WaveOp op = WaveOp::dpp;
        uint32_t in = 0;
        uint32_t old = 0;
        int arg0 = 0, arg1 = 0, arg2 = 0, arg3 = 0;
        uint32_t out = 0;
        uint64_t out64 = 0;
        double vin[6] = {};
        double vout[4] = {};
    };

    struct Dim
    {
        unsigned x = 1, y = 1, z = 1;
    };

    class Workgroup;
    Workgroup *current_group();
    Thread *current_thread();

    [[noreturn]] inline void fail(const std::string &msg)
    {
        std::fprintf(stderr, "SIMT MODEL ERROR: %s\n", msg.c_str());
        std::fflush(stderr);
        std::abort();
    }

    class Workgroup
    {
    public:
        Workgroup(int threads, Dim block_idx, Dim grid_dim, std::function<void()> body, size_t stack_bytes)
            : threads_(threads), body_(std::move(body)), block_idx_(block_idx), grid_dim_(grid_dim)
        {
            if (threads % kWave != 0)
            {
                fail("workgroup size must be a multiple of 64");
            }
            t_.resize(threads);
            for (int i = 0; i < threads; ++i)
            {
                Thread &th = t_[i];
                th.tid = i;
                th.stack.reset(new unsigned char[stack_bytes]);
                // initial frame: six zero registers, then the return address of simt_switch = entry; the stack
                // pointer at entry's first instruction is 8 mod 16, as after a call
                uintptr_t top = reinterpret_cast<uintptr_t>(th.stack.get() + stack_bytes) & ~uintptr_t(15);
                void **sp = reinterpret_cast<void **>(top);
                *--sp = nullptr; // fake return address of entry (never used: entry does not return)
                *--sp = reinterpret_cast<void *>(&Workgroup::entry);
                for (int r = 0; r < 6; ++r)
                {
                    *--sp = nullptr;
                }
                th.sp = sp;
            }
        }

        void run()
        {
            Workgroup *prev = g_group;
            g_group = this;
            for (;;)
            {
                bool ran = false;
                for (int i = 0; i < threads_; ++i)
                {
                    if (t_[i].state == Wait::running)
                    {
                        cur_ = i;
                        simt_switch(&sched_sp_, t_[i].sp);
                        cur_ = -1;
                        ran = true;
                    }
                }
                if (resolve())
                {
                    continue;
                }
                if (!ran)
                {
                    bool all_done = true;
                    for (const Thread &th : t_)
                    {
                        all_done = all_done && th.state == Wait::done;
                    }
                    if (all_done)
                    {
                        break;
                    }
                    report_deadlock();
                }
            }
            g_group = prev;
        }

        Thread &self() { return t_[cur_]; }
        Dim block_idx() const { return block_idx_; }
        Dim grid_dim() const { return grid_dim_; }
        Dim block_dim() const { return Dim{static_cast<unsigned>(threads_), 1, 1}; }
        uint64_t barriers() const { return barriers_; }
        uint64_t wave_ops() const { return wave_ops_; }

        // Called on a fiber: park at a synchronisation point and return to the scheduler.
        void park()
        {
            Thread &th = self();
            simt_switch(&th.sp, sched_sp_);
        }

    private:
        [[noreturn]] static void entry()
        {
            Workgroup *g = g_group;
            g->body_();
            Thread &th = g->self();
            th.state = Wait::done;
            void *dead = nullptr;
            simt_switch(&dead, g->sched_sp_); // never resumed
            std::abort();
        }

        static std::string where(const Thread &th)
        {
            switch (th.state)
            {
            case Wait::running:
                return "running";
            case Wait::done:
                return "left the kernel";
            case Wait::barrier:
                return std::string("barrier at ") + th.file + ":" + std::to_string(th.line);
            case Wait::wave:
                return std::string("wave op at ") + th.file + ":" + std::to_string(th.line);
            }
            return "?";
        }

        [[noreturn]] void report_deadlock()
        {
            std::string msg = "deadlock in workgroup " + std::to_string(block_idx_.x) + ":";
            std::string last;
            int first = 0;
            for (int i = 0; i <= threads_; ++i)
            {
                const std::string w = (i < threads_) ? where(t_[i]) : std::string("<end>");
                if (i == threads_ || (i > 0 && w != last))
                {
                    msg += "\n  threads " + std::to_string(first) + "-" + std::to_string(i - 1) + ": " + last;
                    first = i;
                }
                last = w;
            }
            fail(msg);
        }

        // Releases complete wavefront operations and a complete workgroup barrier; true if anything was released.
        bool resolve()
        {
            bool released = false;
            for (int w = 0; w < threads_ / kWave; ++w)
            {
                Thread *lanes = &t_[w * kWave];
                int at_op = 0;
                for (int l = 0; l < kWave; ++l)
                {
                    at_op += (lanes[l].state == Wait::wave) ? 1 : 0;
                }
                if (at_op == 0)
                {
                    continue;
                }
                if (at_op != kWave)
                {
                    // lanes of one wavefront parked at different places: only an error once nothing else can run
                    bool others_running = false;
                    for (int l = 0; l < kWave; ++l)
                    {
                        others_running = others_running || lanes[l].state == Wait::running;
                    }
                    if (!others_running)
                    {
                        std::string msg = "wavefront " + std::to_string(w) + " of workgroup " +
                                          std::to_string(block_idx_.x) + ": " + std::to_string(at_op) +
                                          " lanes at a wave op, the others elsewhere; lane 0 " + where(lanes[0]) +
                                          ", lane 63 " + where(lanes[kWave - 1]);
                        fail(msg);
                    }
                    continue;
                }
                for (int l = 1; l < kWave; ++l)
                {
                    if (lanes[l].line != lanes[0].line || std::strcmp(lanes[l].file, lanes[0].file) != 0 ||
                        lanes[l].op != lanes[0].op || lanes[l].arg0 != lanes[0].arg0 ||
                        lanes[l].arg1 != lanes[0].arg1 || lanes[l].arg2 != lanes[0].arg2 ||
                        lanes[l].arg3 != lanes[0].arg3)
                    {
                        fail("lanes of wavefront " + std::to_string(w) + " at different wave ops: lane 0 " +
                             where(lanes[0]) + ", lane " + std::to_string(l) + " " + where(lanes[l]));
                    }
                }
                compute_wave(lanes);
                for (int l = 0; l < kWave; ++l)
                {
                    lanes[l].state = Wait::running;
                }
                ++wave_ops_;
                released = true;
            }
            int at_barrier = 0, done = 0;
            for (const Thread &th : t_)
            {
                at_barrier += (th.state == Wait::barrier) ? 1 : 0;
                done += (th.state == Wait::done) ? 1 : 0;
            }
            if (at_barrier > 0 && at_barrier + done == threads_)
            {
                const Thread *ref = nullptr;
                for (const Thread &th : t_)
                {
                    if (th.state == Wait::barrier)
                    {
                        ref = &th;
                        break;
                    }
                }
                if (done > 0)
                {
                    fail("workgroup " + std::to_string(block_idx_.x) + ": " + std::to_string(at_barrier) +
                         " threads wait at " + where(*ref) + " while " + std::to_string(done) +
                         " have left the kernel");
                }
                for (const Thread &th : t_)
                {
                    if (th.line != ref->line || std::strcmp(th.file, ref->file) != 0)
                    {
                        fail("workgroup " + std::to_string(block_idx_.x) + ": threads at different barriers: thread " +
                             std::to_string(ref->tid) + " " + where(*ref) + ", thread " + std::to_string(th.tid) +
                             " " + where(th));
                    }
                }
                for (Thread &th : t_)
                {
                    th.state = Wait::running;
                }
                ++barriers_;
                released = true;
            }
            return released;
        }

        // DPP data movement (GFX9 encodings) and the other cross-lane reads.  A source lane outside the pattern
        // gives 0 with bound_ctrl set and leaves `old` otherwise; rows and banks masked off keep `old`.
        static void compute_wave(Thread *lanes)
        {
            const WaveOp op = lanes[0].op;
            if (op == WaveOp::ballot)
            {
                uint64_t mask = 0;
                for (int l = 0; l < kWave; ++l)
                {
                    mask |= (lanes[l].in != 0) ? (uint64_t(1) << l) : 0;
                }
                for (int l = 0; l < kWave; ++l)
                {
                    lanes[l].out64 = mask;
                }
                return;
            }
            if (op == WaveOp::mfma_f32 || op == WaveOp::mfma_f64)
            {
                // lane l supplies A[l % 16][l / 16] (vin[0]) and B[l / 16][l % 16] (vin[1]) and holds the accumulator
                // entries D[row(l, v)][l % 16], v = 0..3 (vin[2..5] in, vout out): row = 4 (l / 16) + v for FP32,
                // l / 16 + 4 v for FP64.
                for (int l = 0; l < kWave; ++l)
                {
                    const int j = l % 16;
                    for (int v = 0; v < 4; ++v)
                    {
                        const int i = (op == WaveOp::mfma_f32) ? 4 * (l / 16) + v : l / 16 + 4 * v;
                        double acc = lanes[l].vin[2 + v];
                        for (int k = 0; k < 4; ++k)
                        {
                            acc += lanes[i + 16 * k].vin[0] * lanes[j + 16 * k].vin[1];
                        }
                        lanes[l].vout[v] = (op == WaveOp::mfma_f32) ? static_cast<double>(static_cast<float>(acc)) : acc;
                    }
                }
                return;
            }
            for (int l = 0; l < kWave; ++l)
            {
                Thread &d = lanes[l];
                if (op == WaveOp::readlane)
                {
                    const int src = d.arg0;
                    if (src < 0 || src >= kWave)
                    {
                        fail("readlane of lane " + std::to_string(src));
                    }
                    d.out = lanes[src].in;
                    continue;
                }
                if (op == WaveOp::shfl)
                {
                    const int src = static_cast<int>(d.old); // source lane named by this lane
                    if (src < 0 || src >= kWave)
                    {
                        fail("shuffle from lane " + std::to_string(src));
                    }
                    d.out = lanes[src].in;
                    continue;
                }
                const int ctrl = d.arg0, row_mask = d.arg1, bank_mask = d.arg2;
                const bool bound_ctrl = d.arg3 != 0;
                const int row = l / 16, col = l % 16;
                if (!((row_mask >> row) & 1) || !((bank_mask >> (col / 4)) & 1))
                {
                    d.out = d.old;
                    continue;
                }
                int src = -1;
                if (ctrl >= 0x000 && ctrl <= 0x0FF)
                {
                    src = (l & ~3) + ((ctrl >> (2 * (l & 3))) & 3); // quad_perm
                }
                else if (ctrl >= 0x101 && ctrl <= 0x10F)
                {
                    const int k = ctrl - 0x100; // row_shl
                    src = (col + k < 16) ? l + k : -1;
                }
                else if (ctrl >= 0x111 && ctrl <= 0x11F)
                {
                    const int k = ctrl - 0x110; // row_shr
                    src = (col - k >= 0) ? l - k : -1;
                }
                else if (ctrl >= 0x121 && ctrl <= 0x12F)
                {
                    const int k = ctrl - 0x120; // row_ror
                    src = row * 16 + (col - k + 16) % 16;
                }
                else if (ctrl == 0x130)
                {
                    src = (l + 1 < kWave) ? l + 1 : -1; // wave_shl:1
                }
                else if (ctrl == 0x134)
                {
                    src = (l + 1) % kWave; // wave_rol:1
                }
                else if (ctrl == 0x138)
                {
                    src = (l >= 1) ? l - 1 : -1; // wave_shr:1
                }
                else if (ctrl == 0x13C)
                {
                    src = (l + kWave - 1) % kWave; // wave_ror:1
                }
                else if (ctrl == 0x140)
                {
                    src = row * 16 + (15 - col); // row_mirror
                }
                else if (ctrl == 0x141)
                {
                    src = row * 16 + (col & 8) + (7 - (col & 7)); // row_half_mirror
                }
                else if (ctrl == 0x142)
                {
                    src = (row >= 1) ? (row - 1) * 16 + 15 : -1; // row_bcast:15
                }
                else if (ctrl == 0x143)
                {
                    src = (row >= 2) ? 31 : -1; // row_bcast:31
                }
                else
                {
                    fail("unmodelled DPP control " + std::to_string(ctrl));
                }
                if (src < 0)
                {
                    d.out = bound_ctrl ? 0u : d.old;
                }
                else
                {
                    d.out = lanes[src].in;
                }
            }
        }

        static inline Workgroup *g_group = nullptr;
        friend Workgroup *current_group();
        friend Thread *current_thread();

        int threads_;
        std::function<void()> body_;
        Dim block_idx_, grid_dim_;
        std::vector<Thread> t_;
        void *sched_sp_ = nullptr;
        int cur_ = -1;
        uint64_t barriers_ = 0;
        uint64_t wave_ops_ = 0;
    };

    inline Workgroup *current_group() { return Workgroup::g_group; }
    inline Thread *current_thread() { return &Workgroup::g_group->self(); }

    inline void barrier(const char *file, int line)
    {
        Thread *th = current_thread();
        th->state = Wait::barrier;
        th->file = file;
        th->line = line;
        current_group()->park();
    }

    // Matrix multiply-accumulate of a wavefront: returns this lane's four accumulator entries.
    inline void wave_mfma(WaveOp op, double a, double b, const double (&c)[4], double (&d)[4], const char *file, int line)
    {
        Thread *th = current_thread();
        th->state = Wait::wave;
        th->op = op;
        th->in = 0;
        th->old = 0;
        th->arg0 = th->arg1 = th->arg2 = th->arg3 = 0;
        th->vin[0] = a;
        th->vin[1] = b;
        for (int v = 0; v < 4; ++v)
        {
            th->vin[2 + v] = c[v];
        }
        th->file = file;
        th->line = line;
        current_group()->park();
        for (int v = 0; v < 4; ++v)
        {
            d[v] = th->vout[v];
        }
    }

    inline uint64_t wave_ballot(bool pred, const char *file, int line)
    {
        Thread *th = current_thread();
        th->state = Wait::wave;
        th->op = WaveOp::ballot;
        th->in = pred ? 1u : 0u;
        th->old = 0;
        th->arg0 = th->arg1 = th->arg2 = th->arg3 = 0;
        th->file = file;
        th->line = line;
        current_group()->park();
        return th->out64;
    }

    inline uint32_t wave_op(WaveOp op, uint32_t in, uint32_t old, int a0, int a1, int a2, int a3, const char *file,
                            int line)
    {
        Thread *th = current_thread();
        th->state = Wait::wave;
        th->op = op;
        th->in = in;
        th->old = old;
        th->arg0 = a0;
        th->arg1 = a1;
        th->arg2 = a2;
        th->arg3 = a3;
        th->file = file;
        th->line = line;
        current_group()->park();
        return th->out;
    }

    // Runs `grid` workgroups of `threads` threads one after another, in the order given by `order` (block
    // indices; all of 0..grid-1 when empty).  `kernel` is called on every fiber.
    struct LaunchStats
    {
        uint64_t barriers = 0;
        uint64_t wave_ops = 0;
    };

    // Two-dimensional grid: workgroups (x, y) for y in 0..grid_y-1, x in 0..grid_x-1, one after another.
    inline LaunchStats launch2d(int grid_x, int grid_y, int threads, const std::function<void()> &kernel,
                                size_t stack_bytes = 128 * 1024)
    {
        LaunchStats st;
        for (int y = 0; y < grid_y; ++y)
        {
            for (int x = 0; x < grid_x; ++x)
            {
                Workgroup wg(threads, Dim{static_cast<unsigned>(x), static_cast<unsigned>(y), 0},
                             Dim{static_cast<unsigned>(grid_x), static_cast<unsigned>(grid_y), 1}, kernel, stack_bytes);
                wg.run();
                st.barriers += wg.barriers();
                st.wave_ops += wg.wave_ops();
            }
        }
        return st;
    }

    inline LaunchStats launch(int grid, int threads, const std::function<void()> &kernel,
                              std::vector<int> order = {}, size_t stack_bytes = 128 * 1024)
    {
        if (order.empty())
        {
            for (int b = 0; b < grid; ++b)
            {
                order.push_back(b);
            }
        }
        LaunchStats st;
        for (int b : order)
        {
            Workgroup wg(threads, Dim{static_cast<unsigned>(b), 0, 0}, Dim{static_cast<unsigned>(grid), 1, 1}, kernel,
                         stack_bytes);
            wg.run();
            st.barriers += wg.barriers();
            st.wave_ops += wg.wave_ops();
        }
        return st;
    }
} // namespace simt
