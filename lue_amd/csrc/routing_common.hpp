// Shared definitions of the flow-routing kernels: ldd decoding, neighbour tables, tile geometry.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lue_amd {

    // ldd no-data value: numeric_limits<uint8>::max()
    // (reference: source/framework/algorithm/include/lue/framework/algorithm/flow_direction.hpp:7-35)
    constexpr std::uint8_t no_data_ldd = 255;

    // Neighbour k of a cell, clockwise from north: N, NE, E, SE, S, SW, W, NW.
    // idx0 is the row (north = -1), idx1 the column. The offsets + 1 are 2-bit fields of a constant (bit
    // extract and subtract for a k known only at run time, where a chain of comparisons compiles to
    // branches): rows 0 0 1 2 2 2 1 0, columns 1 2 2 2 1 0 0 0.
    __host__ __device__ constexpr int neighbour_dr(int k)
    {
        return static_cast<int>((0x1A90u >> (2 * k)) & 3u) - 1;
    }

    __host__ __device__ constexpr int neighbour_dc(int k)
    {
        return static_cast<int>((0x01A9u >> (2 * k)) & 3u) - 1;
    }

    static_assert(neighbour_dr(0) == -1 && neighbour_dc(0) == 0 && neighbour_dr(1) == -1 && neighbour_dc(1) == 1 &&
                      neighbour_dr(2) == 0 && neighbour_dc(2) == 1 && neighbour_dr(3) == 1 && neighbour_dc(3) == 1 &&
                      neighbour_dr(4) == 1 && neighbour_dc(4) == 0 && neighbour_dr(5) == 1 && neighbour_dc(5) == -1 &&
                      neighbour_dr(6) == 0 && neighbour_dc(6) == -1 && neighbour_dr(7) == -1 && neighbour_dc(7) == -1,
                  "neighbour offsets N, NE, E, SE, S, SW, W, NW");

    // ldd keypad code (7 8 9 / 4 5 6 / 1 2 3) -> neighbour index of the downstream cell, or -1 for a
    // sink (5) and for any code that is not a direction, which LUE also treats as a sink
    // (detail/flow_direction.hpp:11-75 leaves the offsets at zero for such codes).
    // Nibble `code` of the table holds the direction, 0xF for none.
    __host__ __device__ inline int ldd_direction(std::uint32_t code)
    {
        constexpr std::uint64_t table = 0xFFFFFF10'72F6345Full;
        // code: 0  1  2  3  4  5  6  7  8  9
        // dir : F  5  4  3  6  F  2  7  0  1
        int const d = code < 10 ? static_cast<int>((table >> (4 * code)) & 0xFu) : 0xF;
        return d == 0xF ? -1 : d;
    }

    __host__ __device__ constexpr int opposite(int k)
    {
        return (k + 4) & 7;
    }

    // Status codes of the C ABI (include/lue_amd/routing.h).
    enum Status : int {
        status_success = 0,
        status_invalid_argument = 1,
        status_shape_mismatch = 2,
        status_out_of_memory = 3,
        status_hip_error = 4,
        status_invalid_network = 5,
        status_no_device = 6,
    };

    // Pointer into device memory (address space 1). Accesses through it compile to global_* instructions.
    // Through generic pointers the compiler may merge an LDS access and an HBM access of two branches
    // into one flat access, which then waits on both the LDS and the vector-memory counters.
    template<typename P>
    using global_ptr = P __attribute__((address_space(1)))*;

    template<typename P>
    __device__ inline global_ptr<P> as_global(P* p)
    {
        return (global_ptr<P>)p;
    }

    // Operation kinds of the accumulating router.
    enum Op : int
    {
        op_accu = 0,
        op_threshold = 1,
        op_fraction = 2,
        op_kinematic_wave = 3,
        op_capacity = 4,
        op_trigger = 5
    };

    // Operations with two outputs (outflow, remainder) split from the material that enters a cell.
    constexpr bool is_split_op(int op)
    {
        return op == op_threshold || op == op_fraction || op == op_capacity || op == op_trigger;
    }

}  // namespace lue_amd
