// step_erd.hip -- the elongated rhombic dodecahedron (ERD) cell shape.
//
// searchSimulationERD.slang:59-217: cells in the linear nibble layout (cell
// x + X*(y + Y*z), 8 per 32-bit word, no row padding), 12 face neighbours
// whose offsets depend on the parity of the cell's y (odd rows sit half a cell
// further along x and z), 4 edge neighbours, no corner neighbours; the W flag
// wraps every axis.  One lane per packed word, the reference's decomposition:
// this shape is a correctness path, not a measured one.
#include "step_kernels.h"

namespace cave {

// elongatedRhombicDodecahedron.slang:6-27 (the 12 nearest cell centres of the
// honeycomb, then the next 4)
__constant__ int8_t kErdFace[2][12][3] = {
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1},
     {-1, -1, 0}, {-1, -1, -1}, {0, -1, -1}, {-1, 1, 0}, {-1, 1, -1}, {0, 1, -1}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1},
     {1, -1, 0}, {1, -1, 1}, {0, -1, 1}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}},
};
__constant__ int8_t kErdEdge[4][3] = {{1, 0, 1}, {1, 0, -1}, {-1, 0, 1}, {-1, 0, -1}};

struct ErdGeom {
    uint32_t X, Y, Z;
    uint32_t cells;  // X*Y*Z
    uint32_t words;  // ceil(cells / 8)
};

template <bool WRAP>
__device__ __forceinline__ uint32_t erd_alive(const uint32_t *__restrict__ in, const ErdGeom &g, int x, int y, int z,
                                              uint32_t M)
{
    if constexpr (WRAP) {
        x = (x + (int)g.X) % (int)g.X;
        y = (y + (int)g.Y) % (int)g.Y;
        z = (z + (int)g.Z) % (int)g.Z;
    } else {
        if (x < 0 || x >= (int)g.X || y < 0 || y >= (int)g.Y || z < 0 || z >= (int)g.Z) return 0u;
    }
    const uint32_t i = (uint32_t)x + g.X * ((uint32_t)y + g.Y * (uint32_t)z);
    return ((in[i >> 3] >> (4u * (i & 7u))) & 15u) == M ? 1u : 0u;
}

template <bool WRAP>
__global__ void __launch_bounds__(256) k_step_erd(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                  ErdGeom g, DevRule rule, StatPart *__restrict__ partials,
                                                  const uint32_t *__restrict__ done)
{
    __shared__ int32_t red[10 * 4];
    if (done && *done) return;  // the manager already ended this simulation
    StatAcc acc;
    // grid-stride over the words (the launch is capped at kMaxStepBlocks)
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < g.words; w += gridDim.x * blockDim.x) {
        const uint32_t cur_word = in[w];
        uint32_t out_word = 0u;
        const bool fF = rule.flags & 1u, fE = rule.flags & 2u;
        for (uint32_t i = 0; i < 8u; i++) {
            const uint32_t idx = w * 8u + i;
            if (idx >= g.cells) break;
            const int x = (int)(idx % g.X);
            const uint32_t rest = idx / g.X;
            const int y = (int)(rest % g.Y);
            const int z = (int)(rest / g.Y);
            const int cfg = y & 1;
            uint32_t nb = 0u;
            if (fF)
                for (int k = 0; k < 12; k++)
                    nb += erd_alive<WRAP>(in, g, x + kErdFace[cfg][k][0], y + kErdFace[cfg][k][1],
                                          z + kErdFace[cfg][k][2], rule.M);
            if (fE)
                for (int k = 0; k < 4; k++)
                    nb += erd_alive<WRAP>(in, g, x + kErdEdge[k][0], y + kErdEdge[k][1], z + kErdEdge[k][2], rule.M);
            const uint32_t cur = (cur_word >> (4u * i)) & 15u;
            uint32_t ns;
            // searchSimulationERD.slang:161-197
            if (cur == rule.M) {
                if ((rule.surv_lut >> nb) & 1u) { ns = rule.M; acc.surv++; }
                else { ns = cur - 1u; acc.dying++; }
            } else if (cur == 0u) {
                if ((rule.born_lut >> nb) & 1u) { ns = rule.M; acc.born++; }
                else { ns = 0u; acc.dead++; }
            } else {
                ns = cur - 1u;
                acc.dying++;
            }
            out_word |= (ns & 15u) << (4u * i);
            if (ns == rule.M) {
                acc.bmin[0] = min(acc.bmin[0], x); acc.bmax[0] = max(acc.bmax[0], x);
                acc.bmin[1] = min(acc.bmin[1], y); acc.bmax[1] = max(acc.bmax[1], y);
                acc.bmin[2] = min(acc.bmin[2], z); acc.bmax[2] = max(acc.bmax[2], z);
            }
        }
        out[w] = out_word;
    }
    block_reduce_stats(acc, &partials[blockIdx.x], red);
}

// Initial cells in the linear layout (searchSystem.cpp:264-303, linear branch).
__global__ void k_init_erd(uint32_t *cells, ErdGeom g, DevInit init, uint32_t live_nibble)
{
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= g.words) return;
    uint32_t out = 0u;
    for (uint32_t i = 0; i < 8u; i++) {
        const uint32_t idx = w * 8u + i;
        if (idx >= g.cells) break;
        const uint32_t x = idx % g.X, rest = idx / g.X;
        if (init_alive(init, x, rest % g.Y, rest / g.Y)) out |= live_nibble << (4u * i);
    }
    cells[w] = out;
}

static ErdGeom erd_geom(uint32_t X, uint32_t Y, uint32_t Z)
{
    ErdGeom g;
    g.X = X;
    g.Y = Y;
    g.Z = Z;
    g.cells = X * Y * Z;
    g.words = (g.cells + 7u) / 8u;
    return g;
}

uint32_t erd_blocks(uint32_t X, uint32_t Y, uint32_t Z)
{
    return std::min<uint32_t>((erd_geom(X, Y, Z).words + 255u) / 256u, kMaxStepBlocks);
}

int launch_erd(const uint32_t *in, uint32_t *out, uint32_t X, uint32_t Y, uint32_t Z, const DevRule &rule, bool wrap,
               StatPart *partials, hipStream_t st, const uint32_t *done)
{
    const ErdGeom g = erd_geom(X, Y, Z);
    const dim3 grid(erd_blocks(X, Y, Z)), block(256);
    if (wrap) hipLaunchKernelGGL(k_step_erd<true>, grid, block, 0, st, in, out, g, rule, partials, done);
    else hipLaunchKernelGGL(k_step_erd<false>, grid, block, 0, st, in, out, g, rule, partials, done);
    CAVE_HIP(hipGetLastError());
    return CAVE_OK;
}

int launch_init_erd(uint32_t *cells, uint32_t X, uint32_t Y, uint32_t Z, const DevInit &init, uint32_t M,
                    hipStream_t st)
{
    const ErdGeom g = erd_geom(X, Y, Z);
    hipLaunchKernelGGL(k_init_erd, dim3((g.words + 255u) / 256u), dim3(256), 0, st, cells, g, init, M & 15u);
    CAVE_HIP(hipGetLastError());
    return CAVE_OK;
}

}  // namespace cave
