"""Independent brute-force models of the routing operations, for tiny rasters.

These do not restate LUE's router: they use other algorithms (walk from every cell to its outlet;
Kahn's algorithm over the cell graph; closed forms) so that a plausible mistake in the oracle or in the
GPU path (a dropped term, a wrong neighbour offset, a transposed index, a wrong split) shows up as a
difference. Pure Python loops: only for small inputs.
"""
from __future__ import annotations

import math

import numpy as np

OFFSETS = {7: (-1, -1), 8: (-1, 0), 9: (-1, 1), 4: (0, -1), 6: (0, 1), 1: (1, -1), 2: (1, 0), 3: (1, 1)}
ND = 255


def downstream(ldd, r, c):
    """Downstream cell of (r, c) or None for sinks, flow off the array and flow into no-data."""
    code = int(ldd[r, c])
    if code not in OFFSETS:
        return None
    dr, dc = OFFSETS[code]
    r2, c2 = r + dr, c + dc
    rows, cols = ldd.shape
    if not (0 <= r2 < rows and 0 <= c2 < cols) or ldd[r2, c2] == ND:
        return None
    return r2, c2


def random_ldd(rng, rows, cols, nodata_fraction=0.0, pit_fraction=0.0, flat_fraction=0.0):
    """Acyclic ldd: steepest descent on a random surface; optional no-data, pits and random codes on
    'flat' cells that still point downhill (a cell only points to a strictly lower neighbour)."""
    z = rng.random((rows, cols))
    ldd = np.full((rows, cols), 5, dtype=np.uint8)
    code_of = {v: k for k, v in OFFSETS.items()}
    for r in range(rows):
        for c in range(cols):
            best, best_drop = None, 0.0
            lower = []
            for (dr, dc) in code_of:
                r2, c2 = r + dr, c + dc
                if 0 <= r2 < rows and 0 <= c2 < cols:
                    drop = (z[r, c] - z[r2, c2]) / (math.sqrt(2.0) if dr and dc else 1.0)
                    if drop > 0:
                        lower.append((dr, dc))
                    if drop > best_drop:
                        best, best_drop = (dr, dc), drop
            if best is None:
                # no lower neighbour: a pit, or flow off the array edge when on the border
                if r == 0:
                    ldd[r, c] = 8
                elif r == rows - 1:
                    ldd[r, c] = 2
                elif c == 0:
                    ldd[r, c] = 4
                elif c == cols - 1:
                    ldd[r, c] = 6
                else:
                    ldd[r, c] = 5
                continue
            if lower and rng.random() < flat_fraction:
                best = lower[rng.integers(len(lower))]
            ldd[r, c] = code_of[best]
            if rng.random() < pit_fraction:
                ldd[r, c] = 5
    if nodata_fraction > 0:
        ldd[rng.random((rows, cols)) < nodata_fraction] = ND
    return ldd


def inflow_count(ldd):
    rows, cols = ldd.shape
    out = np.zeros((rows, cols), dtype=np.int64)
    for r in range(rows):
        for c in range(cols):
            if ldd[r, c] == ND:
                continue
            d = downstream(ldd, r, c)
            if d is not None:
                out[d] += 1
    out[ldd == ND] = 255
    return out.astype(np.uint8)


def accu_by_walking(ldd, material):
    """Every cell adds its material to each cell on its path to the outlet (O(N * L))."""
    rows, cols = ldd.shape
    out = np.zeros((rows, cols), dtype=np.float64)
    nan_mark = np.zeros((rows, cols), dtype=bool)
    for r in range(rows):
        for c in range(cols):
            if ldd[r, c] == ND:
                continue
            m = float(material[r, c]) if np.ndim(material) else float(material)
            bad = math.isnan(m) or m < 0
            cell = (r, c)
            while cell is not None:
                if bad:
                    nan_mark[cell] = True
                else:
                    out[cell] += m
                cell = downstream(ldd, *cell)
    out[nan_mark] = np.nan
    out[ldd == ND] = np.nan
    return out


def topological_order(ldd):
    """Kahn's algorithm over cells (not LUE's walk order)."""
    rows, cols = ldd.shape
    indeg = np.zeros((rows, cols), dtype=np.int64)
    for r in range(rows):
        for c in range(cols):
            if ldd[r, c] != ND:
                d = downstream(ldd, r, c)
                if d is not None:
                    indeg[d] += 1
    ready = [(r, c) for r in range(rows) for c in range(cols) if ldd[r, c] != ND and indeg[r, c] == 0]
    order = []
    while ready:
        cell = ready.pop()
        order.append(cell)
        d = downstream(ldd, *cell)
        if d is not None:
            indeg[d] -= 1
            if indeg[d] == 0:
                ready.append(d)
    return order


def _value(x, r, c):
    return float(x[r, c]) if np.ndim(x) else float(x)


def route_split(ldd, material, param, kind):
    """accu_threshold / accu_fraction / accu_capacity / accu_trigger over a topological order; returns
    (outflow, remainder)."""
    rows, cols = ldd.shape
    inflow = np.zeros((rows, cols))
    out = np.full((rows, cols), np.nan)
    rem = np.full((rows, cols), np.nan)
    poisoned = np.zeros((rows, cols), dtype=bool)
    for (r, c) in topological_order(ldd):
        m = _value(material, r, c)
        p = _value(param, r, c)
        valid = not (math.isnan(m) or math.isnan(p) or m < 0 or p < 0 or (kind == "fraction" and p > 1))
        if poisoned[r, c] or not valid:
            poisoned[r, c] = True
        else:
            total = inflow[r, c] + m
            if kind == "threshold":
                o, rm = (total - p, p) if total > p else (0.0, total)
            elif kind == "capacity":
                o, rm = (p, total - p) if total > p else (total, 0.0)
            elif kind == "trigger":
                o, rm = (total, 0.0) if not total < p else (0.0, total)
            else:
                o, rm = p * total, total - p * total
            out[r, c], rem[r, c] = o, rm
        d = downstream(ldd, r, c)
        if d is not None:
            if poisoned[r, c]:
                poisoned[d] = True
            else:
                inflow[d] += out[r, c]
    return out, rem


def kinematic_wave_linear(ldd, q_cur, q_lat, alpha, dt, dx):
    """kinematic_wave with beta = 1: the equation (dt/dx) Q + alpha Q = (dt/dx) Q_up + alpha Q_cur + dt q
    is linear, so Q = K / (dt/dx + alpha). Topological order, no Newton iteration."""
    rows, cols = ldd.shape
    up = np.zeros((rows, cols))
    out = np.full((rows, cols), np.nan)
    for (r, c) in topological_order(ldd):
        a = _value(alpha, r, c)
        k = (dt / dx) * up[r, c] + a * float(q_cur[r, c]) + dt * max(float(q_lat[r, c]), 0.0)
        out[r, c] = k / (dt / dx + a) if (up[r, c] + q_cur[r, c] > 0 or q_lat[r, c] > 0) else 0.0
        d = downstream(ldd, r, c)
        if d is not None:
            up[d] += out[r, c]
    return out


def kinematic_residual(q_new, q_up, q_cur, q_lat, alpha, beta, dt, dx):
    """f(Q) of the kinematic-wave equation, relative to its known terms."""
    known = (dt / dx) * q_up + alpha * q_cur ** beta + dt * q_lat
    return ((dt / dx) * q_new + alpha * q_new ** beta - known) / np.maximum(known, 1e-300)


def d8_flow_direction(elevation):
    """LUE's D8 flow direction, vectorised: pad with numeric_limits<T>::min(), then the eight raw drops in the
    order N, W, E, S, NW, NE, SW, SE, each taken where it exceeds the best so far (start: sink, drop 0);
    NaN (no-data) neighbours are skipped, a NaN centre gives 255."""
    elevation = np.asarray(elevation)
    t = elevation.dtype
    padded = np.pad(elevation, 1, constant_values=np.finfo(t).tiny)
    rows, cols = elevation.shape
    centre = padded[1:-1, 1:-1]
    out = np.full((rows, cols), 5, dtype=np.uint8)
    best = np.zeros((rows, cols), dtype=t)
    for dr, dc, code in ((-1, 0, 8), (0, -1, 4), (0, 1, 6), (1, 0, 2), (-1, -1, 7), (-1, 1, 9), (1, -1, 1),
                         (1, 1, 3)):
        neighbour = padded[1 + dr:1 + dr + rows, 1 + dc:1 + dc + cols]
        with np.errstate(invalid="ignore"):
            drop = (centre - neighbour).astype(t)
            take = ~np.isnan(neighbour) & (drop > best)
        out[take] = code
        best = np.where(take, drop, best)
    out[np.isnan(centre)] = 255
    return out


def material_no_data(dtype):
    """No-data of a material element: NaN for floating point, the minimum of a signed and the maximum of an
    unsigned integer type."""
    dtype = np.dtype(dtype)
    if dtype.kind == "f":
        return dtype.type(np.nan)
    info = np.iinfo(dtype)
    return dtype.type(info.min if dtype.kind == "i" else info.max)


def _is_no_data(a):
    return np.isnan(a) if a.dtype.kind == "f" else a == material_no_data(a.dtype)


def upstream_by_pushing(ldd, material):
    """Each cell pushes its material to its downstream neighbour (vectorised, one direction at a time, in the
    order of the receiving cell's window: from NW, N, NE, W, E, SW, S, SE); a cell is no-data where its ldd or
    material is, or where a no-data material arrives."""
    ldd = np.asarray(ldd)
    material = np.asarray(material)
    rows, cols = material.shape
    total = np.zeros_like(material)
    poisoned = np.zeros(material.shape, dtype=bool)
    source_nd = _is_no_data(material)
    for code in (3, 2, 1, 6, 4, 9, 8, 7):
        dr, dc = OFFSETS[code]
        src = np.zeros((rows + 2, cols + 2), dtype=bool)
        src[1:-1, 1:-1] = ldd == code
        # cells (r, c) whose neighbour (r - dr, c - dc) drains into them
        sent = src[1 - dr:1 - dr + rows, 1 - dc:1 - dc + cols]
        value = np.zeros((rows + 2, cols + 2), dtype=material.dtype)
        value[1:-1, 1:-1] = material
        arriving = value[1 - dr:1 - dr + rows, 1 - dc:1 - dc + cols]
        nd = np.zeros((rows + 2, cols + 2), dtype=bool)
        nd[1:-1, 1:-1] = source_nd
        poisoned |= sent & nd[1 - dr:1 - dr + rows, 1 - dc:1 - dc + cols]
        with np.errstate(over="ignore", invalid="ignore"):
            total = np.where(sent & ~poisoned, total + arriving, total)
    out = total.astype(material.dtype)
    out[poisoned | (ldd == ND) | source_nd] = material_no_data(material.dtype)
    return out


def downstream_by_pulling(ldd, material):
    """Each cell takes the material of its downstream neighbour (0 beyond the array's edge), its own for a pit
    or an unknown code; no-data where its ldd or material is, or where the neighbour's material is."""
    ldd = np.asarray(ldd)
    material = np.asarray(material)
    rows, cols = material.shape
    padded = np.zeros((rows + 2, cols + 2), dtype=material.dtype)
    padded[1:-1, 1:-1] = material
    out = material.copy()
    for code, (dr, dc) in OFFSETS.items():
        target = padded[1 + dr:1 + dr + rows, 1 + dc:1 + dc + cols]
        out = np.where(ldd == code, target, out)
    out[_is_no_data(out) | (ldd == ND) | _is_no_data(material)] = material_no_data(material.dtype)
    return out


def partition_index(extent, part):
    """Partition of each row (or column): extent // part partitions, the last one absorbs the remainder
    (source/framework/core/include/lue/framework/core/domain_decomposition.hpp:91-183)."""
    return np.minimum(np.arange(extent) // part, extent // part - 1)


def accu_info_by_rule(ldd, partition_shape):
    """accu_info's cell classes from each cell's place in the network and in the partition grid, without
    replaying the router's events (definition/accu_info.hpp names the classes):

      - no upstream neighbour: ridge (11);
      - an upstream neighbour in another partition: partition input (14), or partition input and output
        (15) when the downstream cell is outside the partition or the array;
      - otherwise sink (17), partition output (13), or a stream cell: inter-partition (16) when an edge that
        crosses partitions lies upstream, else intra-partition (12).

    A partition input cell that also has an upstream neighbour in its own partition with a crossing edge
    upstream of it gets 14/15 when the message arrives after that neighbour's walk and 17/13/16 (the rule for
    a cell without upstream neighbours in other partitions) otherwise; LUE reads the channels concurrently, so
    either can happen. Returns (classes, alternative), alternative = that second class for such cells, 0
    elsewhere. Pure Python loops: small rasters only.
    """
    ldd = np.asarray(ldd)
    rows, cols = ldd.shape
    prow = partition_index(rows, partition_shape[0])
    pcol = partition_index(cols, partition_shape[1])

    def part(r, c):
        return prow[r], pcol[c]

    def target(r, c):
        """(r2, c2, in_array) of the downstream cell, None for a sink or an unknown code."""
        code = int(ldd[r, c])
        if code not in OFFSETS:
            return None
        dr, dc = OFFSETS[code]
        r2, c2 = r + dr, c + dc
        return r2, c2, 0 <= r2 < rows and 0 <= c2 < cols

    # pending: an edge that crosses partitions lies upstream (mark from each crossing edge down to the outlet)
    pending = np.zeros((rows, cols), dtype=bool)
    upstream = [[[] for _ in range(cols)] for _ in range(rows)]
    for r in range(rows):
        for c in range(cols):
            if ldd[r, c] == ND:
                continue
            t = target(r, c)
            if t is None or not t[2] or ldd[t[0], t[1]] == ND:
                continue
            r2, c2 = t[0], t[1]
            upstream[r2][c2].append((r, c))
            if part(r, c) != part(r2, c2):
                cell = (r2, c2)
                while cell is not None and not pending[cell]:
                    pending[cell] = True
                    t2 = target(*cell)
                    cell = (t2[0], t2[1]) if (t2 is not None and t2[2] and ldd[t2[0], t2[1]] != ND) else None

    classes = np.full((rows, cols), 255, dtype=np.uint8)
    alternative = np.zeros((rows, cols), dtype=np.uint8)
    for r in range(rows):
        for c in range(cols):
            if ldd[r, c] == ND:
                continue
            t = target(r, c)
            sink = t is None
            out = not sink and (not t[2] or part(t[0], t[1]) != part(r, c))
            if sink:
                plain = 17
            elif out:
                plain = 13
            else:
                plain = 16 if pending[r, c] else 12
            ups = upstream[r][c]
            ext = [u for u in ups if part(*u) != part(r, c)]
            if not ups:
                classes[r, c] = 11
            elif ext:
                classes[r, c] = 15 if out else 14
                if any(pending[u] for u in ups if part(*u) == part(r, c)):
                    alternative[r, c] = plain
            else:
                classes[r, c] = plain
    return classes, alternative
