// block.hpp -- block one-sided Jacobi (mixed-precision path): geometry and launchers (block.hip).
#pragma once

#include "common.hpp"

namespace jsvd
{
    constexpr int kB = 32;      // default block width (columns); 16 and 64 are the other instantiated widths
    constexpr int kH = 128;     // row-slab unit: lda and ldv are multiples of it, Gram partials are per kH rows

    // True for the instantiated block widths.
    constexpr bool valid_block_width(int kb) { return kb == 16 || kb == 32 || kb == 64; }

    // Padded geometry of one matrix for the block kernels: lda / ldv are multiples of kH, the column count
    // is padded to nb * kb with zero columns.
    struct BlockGeometry
    {
        int64_t m = 0;      // rows of A
        int64_t n = 0;      // columns of A
        int kb = kB;        // block width (16, 32 or 64); a block pair has 2 kb columns
        int64_t n_pad = 0;  // nb * kb
        int nb = 0;         // blocks
        int Nb = 0;         // nb rounded up to even (round-robin circle size)
        int64_t lda = 0;    // rows of the padded A (multiple of kH)
        int64_t ldv = 0;    // rows of the padded V (multiple of kH, >= n_pad)
        int slabs_a = 0;    // lda / kH
        int slabs_v = 0;    // ldv / kH
        int apply_wgs = 512; // target workgroups of one apply launch (two per CU); the autotuner may change it
        // Inner ordering of a block round at kb = 32: true = only the kb x kb cross pairs of the block pair (kb
        // inner rounds), with one intra-block step per sweep (launch_block_intra) for the pairs inside each block;
        // false = one full cyclic sweep over the block pair's 2 kb columns (2 kb - 1 inner rounds).
        bool cross = false;

        static BlockGeometry make(int64_t m, int64_t n, int kb = kB)
        {
            BlockGeometry g;
            g.m = m;
            g.n = n;
            g.kb = kb;
            g.nb = static_cast<int>((n + kb - 1) / kb);
            g.n_pad = static_cast<int64_t>(g.nb) * kb;
            g.Nb = g.nb + (g.nb & 1);
            g.lda = round_up(m, kH);
            g.ldv = round_up(g.n_pad, kH);
            g.slabs_a = static_cast<int>(g.lda / kH);
            g.slabs_v = static_cast<int>(g.ldv / kH);
            return g;
        }
        [[nodiscard]] int kw() const { return 2 * kb; }
        [[nodiscard]] int rounds() const { return Nb - 1; }
        [[nodiscard]] int pairs() const { return Nb / 2; }
        // Elements of the partial-Gram buffer: at most one kw x kw partial per kH-row slab and pair.
        [[nodiscard]] size_t partial_elems() const
        {
            return static_cast<size_t>(pairs()) * static_cast<size_t>(slabs_a) * kw() * kw();
        }
        // Elements of the rotation buffer: one kw x kw matrix per pair.
        [[nodiscard]] size_t q_elems() const { return static_cast<size_t>(pairs()) * kw() * kw(); }
        // Entries of the per-pair apply flags: a pair per outer-round slot, or a block per intra-step slot.
        [[nodiscard]] size_t flag_elems() const { return static_cast<size_t>(Nb); }
    };

    // Block width for an n-column problem: `requested` (0 = the default kB), halved while it leaves fewer than
    // two blocks (n <= width) and it is above 16.
    inline int effective_block_width(int requested, int64_t n)
    {
        int kb = (requested == 0) ? kB : requested;
        while (kb > 16 && n <= kb)
        {
            kb /= 2;
        }
        return kb;
    }

    // One block round (Gram partials, reduction + inner sweep, apply to A and V) at block width g.kb.  partial
    // holds g.partial_elems() elements, q g.q_elems(), apply_flag pairs ints.
    // perm maps the sweep's logical column order to physical columns (A and V columns are addressed
    // through it); norm2 receives the squared norm of every column after its last rotation of the sweep.
    template <typename T>
    void launch_block_round(const BlockGeometry &g, T *a, T *v, int round, double eps, T *partial, T *q,
                            int *apply_flag, int *rot_count, unsigned long long *offmax, const int *perm,
                            double *norm2, const int *skip_if_zero, hipStream_t stream);

    // The intra-block step of a sweep in the cross ordering (g.cross, kb = 32): every block swept by itself (one
    // cyclic sweep over its kb columns, all blocks at once), with the buffers of launch_block_round (partial and
    // q hold the intra step's smaller matrices; apply_flag needs g.flag_elems() entries).  No-op when !g.cross.
    template <typename T>
    void launch_block_intra(const BlockGeometry &g, T *a, T *v, double eps, T *partial, T *q, int *apply_flag,
                            int *rot_count, unsigned long long *offmax, const int *perm, double *norm2,
                            const int *skip_if_zero, hipStream_t stream);

    // The block round's inner step alone, on `pairs` independent kw x kw Gram matrices (kw = 2 kb; gram[p]
    // row-major, upper triangle read): one cyclic Jacobi sweep in Gram space; q_out[p] (row-major) receives the
    // accumulated rotation and rotations[p] the number of rotated column pairs.
    // cross = true: the cross ordering of the block rounds at kb = 32 (only the kb x kb pairs across the two
    // blocks, kb inner rounds).
    template <typename T>
    void launch_block_eigen(const T *gram, int kb, int pairs, double eps, T *q_out, int *rotations, bool cross,
                            hipStream_t stream);

    // Largest padded column count the one-workgroup ordering kernel sorts (LDS bound).
    constexpr int kSortMax = 8192;

    // perm <- physical columns by descending norm2 (no-op when *skip_if_zero == 0 or n_pad > kSortMax).
    void launch_column_order(const double *norm2, int64_t n_pad, int *perm, const int *skip_if_zero,
                             hipStream_t stream);
    void launch_iota(int *perm, int64_t n, hipStream_t stream);
    // *cont = (*rot > stop): the flag the next sweep's kernels exit on (0 = the phase has ended).
    void launch_sweep_continue(const int *rot, int stop, int *cont, hipStream_t stream);

    void launch_convert_f64_to_f32(const double *src, float *dst, int64_t count, hipStream_t stream);
    void launch_convert_f32_to_f64(const float *src, double *dst, int64_t count, hipStream_t stream);
    void launch_pad_copy_f64(const double *src, int64_t lds, int64_t rows, int64_t cols, double *dst, int64_t ldd,
                             int64_t dcols, bool identity_pad, hipStream_t stream);
    void launch_identity_f32(float *v, int64_t ldv, int64_t cols, hipStream_t stream);
} // namespace jsvd
