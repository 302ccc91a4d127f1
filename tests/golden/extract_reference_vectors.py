#!/usr/bin/env python3
"""Write LUE's own routing test vectors as JSON fixtures (inputs and expected outputs).

This script READS the reference's Boost.Test sources (C++ literals of partitioned arrays) and writes
the data they hold: every input raster, scalar and expected output raster of each test case, as plain
JSON next to this script. It copies no source text: the fixtures hold numbers only. Run it in a
container where the reference is mounted:

    python tests/golden/extract_reference_vectors.py /root/reference

Files read (relative to the reference root):
    source/framework/algorithm/test/flow_accumulation.hpp      (default 9x9 array / 3x3 partitions, ldd names)
    source/framework/algorithm/test/flow_accumulation.cpp.in   (ldd fixtures: merging_streams, spiral_in, ...)
    source/framework/algorithm/test/accu_test.cpp
    source/framework/algorithm/test/accu_threshold_test.cpp
    source/framework/algorithm/test/accu_fraction_test.cpp
    source/framework/algorithm/test/accu_capacity_test.cpp
    source/framework/algorithm/test/accu_trigger_test.cpp
    source/framework/algorithm/test/kinematic_wave_test.cpp
    source/framework/algorithm/test/inflow_count_test.cpp
    source/framework/algorithm/test/d8_flow_direction_test.cpp
    source/framework/algorithm/test/upstream_test.cpp
    source/framework/algorithm/test/downstream_test.cpp
    source/framework/algorithm/test/downstream_distance_test.cpp
    source/framework/algorithm/test/accu_info_test.cpp

A partitioned literal lists one initializer per partition, partitions in row-major partition order
(source/framework/test/include/lue/framework/test/array.hpp:76-105), each partition's cells row-major,
partition shapes per the last-partition-absorbs-the-remainder rule
(source/framework/core/include/lue/framework/core/domain_decomposition.hpp:91-183).
"""
from __future__ import annotations

import json
import math
import os
import re
import sys

LDD_NAMES = {"n": 8, "ne": 9, "e": 6, "se": 3, "s": 2, "sw": 1, "w": 4, "nw": 7, "p": 5, "nd": 255}
# lue::north<FlowDirection> ... (source/framework/algorithm/include/lue/framework/algorithm/flow_direction.hpp)
DIRECTION_NAMES = {"north": 8, "north_east": 9, "east": 6, "south_east": 3, "south": 2, "south_west": 1, "west": 4,
                   "north_west": 7, "sink": 5}
# lue::ridge_cell ... (source/framework/algorithm/include/lue/framework/algorithm/accu_info.hpp:27-63)
CELL_CLASS_NAMES = {"ridge_cell": 11, "intra_partition_stream_cell": 12, "partition_output_cell": 13,
                    "partition_input_cell": 14, "partition_input_output_cell": 15,
                    "inter_partition_stream_cell": 16, "sink_cell": 17, "confluence_cell": 18}
NAN = float("nan")

TOKEN_RE = re.compile(r"""
    (?P<num>-?\d+\.\d*(?:[eE][-+]?\d+)?[fF]?|-?\d*\.\d+(?:[eE][-+]?\d+)?[fF]?|-?\d+(?:[eE][-+]?\d+)?[uUlLfF]*)
  | (?P<id>[A-Za-z_][A-Za-z_0-9]*(?:::[A-Za-z_][A-Za-z_0-9]*)*)
  | (?P<punct>::|[{}()\[\]<>,;=&.*+-])
""", re.VERBOSE)


def strip_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    text = re.sub(r"\{%.*?%\}", " ", text)
    return text


def tokenize(text: str):
    toks = []
    for m in TOKEN_RE.finditer(text):
        kind = m.lastgroup
        toks.append((kind, m.group(kind)))
    return toks


def partition_shapes(array_shape, partition_shape):
    rows, cols = array_shape
    pr, pc = partition_shape
    npr, npc = rows // pr, cols // pc
    heights = [pr] * npr
    widths = [pc] * npc
    heights[-1] += rows - npr * pr
    widths[-1] += cols - npc * pc
    return heights, widths


def assemble(array_shape, partition_shape, partitions):
    heights, widths = partition_shapes(array_shape, partition_shape)
    if len(partitions) != len(heights) * len(widths):
        raise ValueError(f"{len(partitions)} partitions for grid {len(heights)}x{len(widths)}")
    rows, cols = array_shape
    raster = [[None] * cols for _ in range(rows)]
    k = 0
    r0 = 0
    for h in heights:
        c0 = 0
        for w in widths:
            cells = partitions[k]
            if len(cells) != h * w:
                raise ValueError(f"partition {k}: {len(cells)} values for {h}x{w}")
            for i in range(h):
                for j in range(w):
                    raster[r0 + i][c0 + j] = cells[i * w + j]
            c0 += w
            k += 1
        r0 += h
    return raster


class Raster:
    def __init__(self, array_shape, partition_shape, data):
        self.array_shape = tuple(array_shape)
        self.partition_shape = tuple(partition_shape)
        self.data = data  # row-major nested list


class Parser:
    """Evaluates the small expression language the tests use to build partitioned arrays."""

    def __init__(self, toks, env):
        self.toks = toks
        self.i = 0
        self.env = env

    def peek(self, k=0):
        j = self.i + k
        return self.toks[j] if j < len(self.toks) else (None, None)

    def take(self, value=None):
        tok = self.toks[self.i]
        if value is not None and tok[1] != value:
            raise SyntaxError(f"expected {value!r}, got {tok[1]!r} near {self.context()}")
        self.i += 1
        return tok

    def context(self):
        return " ".join(t[1] for t in self.toks[max(0, self.i - 8):self.i + 8])

    def skip_template_args(self):
        if self.peek()[1] != "<":
            return ""
        depth = 0
        parts = []
        while True:
            kind, val = self.take()
            parts.append(val)
            if val == "<":
                depth += 1
            elif val == ">":
                depth -= 1
                if depth == 0:
                    return " ".join(parts)

    def expr(self):
        kind, val = self.peek()
        if val == "{":
            return self.brace_list()
        if val == "-" and self.peek(1)[0] == "num":
            self.take()
            return -self.number(self.take()[1])
        if kind == "num":
            return self.number(self.take()[1])
        if kind == "id":
            return self.identifier_expr()
        raise SyntaxError(f"unexpected {val!r} near {self.context()}")

    @staticmethod
    def number(text):
        text = text.rstrip("uUlLfF") if not re.search(r"[eE.]", text) else text.rstrip("fF")
        return float(text)

    def brace_list(self):
        self.take("{")
        items = []
        while self.peek()[1] != "}":
            items.append(self.expr())
            if self.peek()[1] == ",":
                self.take()
        self.take("}")
        return items

    def call_args(self):
        self.take("(")
        args = []
        while self.peek()[1] != ")":
            args.append(self.expr())
            if self.peek()[1] == ",":
                self.take()
        self.take(")")
        return args

    def identifier_expr(self):
        kind, name = self.take()
        targs = self.skip_template_args()
        while self.peek()[1] == "::":  # e.g. std::numeric_limits<T>::min
            self.take("::")
            name = name + "::" + self.take()[1]
        short = name.split("::")[-1]
        if short == "no_data_value":
            return 255.0 if ("FlowDirection" in targs or "Count" in targs or "CellClass" in targs) else NAN
        if short == "min" and "numeric_limits" in name:
            self.call_args()
            return sys.float_info.min
        if self.peek()[1] == "{" and self.peek(1)[1] == "}":  # policies object
            self.take("{")
            self.take("}")
            return ("policies", name)
        if self.peek()[1] == "(":
            args = self.call_args()
            return self.apply(name, short, targs, args)
        if self.peek()[1] == ".":  # flow_direction.shape()
            self.take(".")
            member = self.take()[1]
            self.call_args()
            obj = self.env[short]
            if member == "shape":
                return list(obj.array_shape)
            raise SyntaxError(member)
        if short in self.env:
            return self.env[short]
        if name.startswith("lue::") and short in CELL_CLASS_NAMES:
            return float(CELL_CLASS_NAMES[short])
        if name.startswith("lue::") and short in DIRECTION_NAMES:
            return float(DIRECTION_NAMES[short])
        if short in LDD_NAMES:
            return float(LDD_NAMES[short])
        raise KeyError(f"unknown identifier {name} near {self.context()}")

    def apply(self, name, short, targs, args):
        if short == "create_partitioned_array":
            array_shape, partition_shape, content = args
            array_shape = [int(v) for v in array_shape]
            partition_shape = [int(v) for v in partition_shape]
            if isinstance(content, list):
                data = assemble(array_shape, partition_shape, content)
            else:
                data = [[content] * array_shape[1] for _ in range(array_shape[0])]
            return Raster(array_shape, partition_shape, data)
        if short in FIXTURES:
            return FIXTURES[short](*args)
        if short == "ones":
            return FIXTURES["filled"](1.0)
        if short == "zeros":
            return FIXTURES["filled"](0.0)
        if short == "no_data":
            return FIXTURES["filled"](255.0 if ("FlowDirection" in targs or "Count" in targs) else NAN)
        if name.startswith("lue::value_policies::") or name.startswith("lue::detail::iterate_to_new_discharge"):
            return ("call", short, args)
        raise KeyError(f"unknown call {name}")


FIXTURES = {}


def split_statements(toks):
    stmts, cur, depth = [], [], 0
    for tok in toks:
        if tok[1] in "{([":
            depth += 1
        elif tok[1] in "})]":
            depth -= 1
        if tok[1] == ";" and depth == 0:
            stmts.append(cur)
            cur = []
        else:
            cur.append(tok)
            if tok[1] == "}" and depth == 0 and cur[0][1] == "{":
                stmts.append(cur)  # a nested block scope
                cur = []
    if cur:
        stmts.append(cur)
    return stmts


DECL_RE = re.compile(r"^(?:auto const(?: &)?|auto|lue::Scalar < \w+ > const|Shape const|Shape|FloatingPoint const|"
                     r"lue::ShapeT < [\w:]+ > const|InflowCount|FlowDirection|"
                     r"lue::PartitionedArray < lue::CellClass , 2 >(?: const)?)$")


def run_statements(stmts, env, record):
    for stmt in stmts:
        if not stmt:
            continue
        vals = [t[1] for t in stmt]
        if vals[0] == "{" and vals[-1] == "}":
            run_statements(split_statements(stmt[1:-1]), env, record)
            continue
        if vals[0] == "if" and vals[-1] == "}":  # if constexpr (supported element) { ... }
            run_statements(split_statements(stmt[vals.index("{") + 1:-1]), env, record)
            continue
        # structured binding: auto const [a, b] = expr
        if vals[0] == "auto" and "[" in vals[:4]:
            lb = vals.index("[")
            rb = vals.index("]")
            names = [v for v in vals[lb + 1:rb] if v != ","]
            eq = vals.index("=")
            value = Parser(stmt[eq + 1:], env).expr()
            env[tuple(names)] = value
            record.append((tuple(names), value))
            continue
        # declarations: <type> NAME = expr | <type> NAME{expr}
        for k in range(1, min(len(stmt), 10)):
            head = " ".join(vals[:k])
            if DECL_RE.match(head) and stmt[k][0] == "id" and k + 1 < len(stmt) and vals[k + 1] in ("=", "{"):
                name = vals[k]
                if vals[k + 1] == "=":
                    value = Parser(stmt[k + 2:], env).expr()
                else:
                    inner = Parser(stmt[k + 1:], env).brace_list()
                    value = inner[0] if len(inner) == 1 else inner
                    if isinstance(value, list) and len(value) == 1 and isinstance(value[0], list):
                        value = value[0]
                env[name] = value
                record.append((name, value))
                break
        else:
            if vals[0] in ("test_inflow_count",):
                args = Parser(stmt[1:], env).call_args()
                record.append(("test_inflow_count", args))
            elif vals[0].startswith("lue::test::check_arrays_are_"):
                args = [t for t in vals[2:-1] if t != ","]
                record.append((vals[0].split("_are_")[1], args))
            elif vals[0] in ("using", "namespace", "BOOST_CHECK", "BOOST_TEST", "std::random_device",
                             "std::default_random_engine", "std::uniform_real_distribution", "for",
                             "lue::PartitionedArray", "if", "MaterialElement", "outflow", "InflowCount"):
                continue
    return record


def parse_fixture_file(path):
    text = strip_comments(open(path).read())
    toks = tokenize(text)
    # default shapes from flow_accumulation.hpp: 9x9 arrays of 3x3 partitions
    base_env = {"array_shape": [9, 9], "partition_shape": [3, 3], **{k: float(v) for k, v in LDD_NAMES.items()}}
    # find 'NAME ( ) { ... }' function bodies
    i = 0
    while i < len(toks):
        if toks[i][0] == "id" and i + 2 < len(toks) and toks[i + 1][1] == "(" and toks[i][1] in (
                "merging_streams", "pcraster_example_flow_direction", "all_no_data", "spiral_in"):
            name = toks[i][1]
            j = i + 2
            while toks[j][1] != "{":
                j += 1
            depth, k = 0, j
            while True:
                if toks[k][1] == "{":
                    depth += 1
                elif toks[k][1] == "}":
                    depth -= 1
                    if depth == 0:
                        break
                k += 1
            body = toks[j + 1:k]
            env = dict(base_env)
            result = {}
            for stmt in split_statements(body):
                vals = [t[1] for t in stmt]
                if vals and vals[0] == "return":
                    result["value"] = Parser(stmt[1:], env).expr()
                else:
                    run_statements([stmt], env, [])
            raster = result["value"]
            FIXTURES[name] = (lambda r: (lambda: r))(raster)
            i = k
        i += 1

    def filled(v):
        return Raster((9, 9), (3, 3), [[v] * 9 for _ in range(9)])

    FIXTURES["filled"] = filled


def case_bodies(path):
    text = strip_comments(open(path).read())
    toks = tokenize(text)
    cases = []
    i = 0
    while i < len(toks):
        if toks[i][1] == "BOOST_AUTO_TEST_CASE":
            name = toks[i + 2][1]
            j = i + 4
            while toks[j][1] != "{":
                j += 1
            depth, k = 0, j
            while True:
                if toks[k][1] == "{":
                    depth += 1
                elif toks[k][1] == "}":
                    depth -= 1
                    if depth == 0:
                        break
                k += 1
            cases.append((name, toks[j + 1:k]))
            i = k
        i += 1
    return cases


def to_json_raster(r: Raster):
    def conv(v):
        if isinstance(v, float) and math.isnan(v):
            return "nan"
        return v
    return [[conv(v) for v in row] for row in r.data]


def arg_json(v):
    if isinstance(v, Raster):
        return {"raster": to_json_raster(v)}
    return {"scalar": "nan" if isinstance(v, float) and math.isnan(v) else v}


def extract(ref_root):
    test_dir = os.path.join(ref_root, "source/framework/algorithm/test")
    parse_fixture_file(os.path.join(test_dir, "flow_accumulation.cpp.in"))
    out = {}

    routing_files = {
        "accu": "accu_test.cpp",
        "accu_threshold": "accu_threshold_test.cpp",
        "accu_fraction": "accu_fraction_test.cpp",
        "accu_capacity": "accu_capacity_test.cpp",
        "accu_trigger": "accu_trigger_test.cpp",
        "kinematic_wave": "kinematic_wave_test.cpp",
    }
    for op, fname in routing_files.items():
        path = os.path.join(test_dir, fname)
        cases = []
        for name, body in case_bodies(path):
            if name in ("overloads", "random_input"):
                continue
            env = {"array_shape": [9, 9], "partition_shape": [3, 3], "lue::test::array_shape": [9, 9]}
            env.update({k: float(v) for k, v in LDD_NAMES.items()})
            record = run_statements(split_statements(body), env, [])
            case = {"name": name, "source": f"source/framework/algorithm/test/{fname}"}
            calls = [(k, v) for k, v in record if isinstance(v, tuple) and v and v[0] == "call"]
            if name in ("dry_cell", "crashed_in_pcraster2"):
                # scalar calls of iterate_to_new_discharge: got/want pairs in order
                got = [v for k, v in record if k == "outflow_we_got"]
                want = [v for k, v in record if k == "outflow_we_want"]
                case["kind"] = "iterate_to_new_discharge"
                case["element"] = "double" if name == "crashed_in_pcraster2" else "float"
                case["calls"] = [
                    {"args": g[2], "want": w} for g, w in zip(got, want)]
                case["compare"] = "close_1e-6"
                cases.append(case)
                continue
            if not calls:
                continue
            target, (_, opname, args) = calls[-1]
            ldd = args[0]
            case["kind"] = opname
            case["array_shape"] = list(ldd.array_shape)
            case["partition_shape"] = list(ldd.partition_shape)
            case["ldd"] = to_json_raster(ldd)
            arg_names = {
                "accu": ["inflow"],
                "accu_threshold": ["inflow", "threshold"],
                "accu_fraction": ["inflow", "fraction"],
                "accu_capacity": ["inflow", "capacity"],
                "accu_trigger": ["inflow", "trigger"],
                "kinematic_wave": ["current_outflow", "inflow", "alpha", "beta", "time_step_duration",
                                   "channel_length"],
            }[opname]
            case["args"] = {n: arg_json(a) for n, a in zip(arg_names, args[1:])}
            compares = [(k, v) for k, v in record if k in ("equal", "close")]
            expected = {}
            for mode, (got_name, want_name) in compares:
                key = "outflow" if "outflow" in got_name else "remainder"
                expected[key] = {"raster": to_json_raster(env[want_name]), "compare": mode}
            if not expected:
                # property-only cases (e.g. all outflow > 0)
                case["property"] = "all_positive"
            case["expected"] = expected
            cases.append(case)
        out[op] = cases

    # d8_flow_direction: every (elevation, expected) pair a case checks with check_arrays_are_equal (nested
    # blocks may hold several); cases without a check (barnes_2017: LUE does not match it) are left out
    path = os.path.join(test_dir, "d8_flow_direction_test.cpp")
    cases = []
    for name, body in case_bodies(path):
        env = {}
        record = run_statements(split_statements(body), env, [])
        last = {}
        k = 0
        for key, value in record:
            if key == "equal":
                got_name, want_name = value
                call = last[got_name]
                elevation = call[2][0]
                want = last[want_name]
                k += 1
                cases.append({
                    "name": name if k == 1 else f"{name}_{k}",
                    "source": "source/framework/algorithm/test/d8_flow_direction_test.cpp",
                    "kind": "d8_flow_direction",
                    "array_shape": list(elevation.array_shape),
                    "partition_shape": list(elevation.partition_shape),
                    "elevation": to_json_raster(elevation),
                    "expected": {"flow_direction": {"raster": to_json_raster(want), "compare": "equal"}},
                })
            elif isinstance(key, str):
                last[key] = value
    out["d8_flow_direction"] = cases

    # upstream / downstream: (flow_direction, material, expected) of every check; the material element
    # type is the case's `using MaterialElement = ...` (the no-data value x depends on it: NaN for floating
    # point, the largest value for unsigned integers)
    element_names = {"FloatingPointElement": "float32", "LargestUnsignedIntegralElement": "uint64"}
    for op in ("upstream", "downstream"):
        fname = f"{op}_test.cpp"
        cases = []
        for name, body in case_bodies(os.path.join(test_dir, fname)):
            vals = [t[1] for t in body]
            at = next(i for i in range(len(vals) - 2) if vals[i] == "using" and vals[i + 1] == "MaterialElement")
            element_tokens = vals[at + 3:vals.index(";", at)]
            element = next(element_names[t.split("::")[-1]] for t in element_tokens
                           if t.split("::")[-1] in element_names)
            env = {}
            record = run_statements(split_statements(body), env, [])
            last = {}
            k = 0
            for key, value in record:
                if key == "equal":
                    got_name, want_name = value
                    _, opname, args = last[got_name]
                    assert opname == op
                    flow_direction, material = args
                    k += 1
                    cases.append({
                        "name": name if k == 1 else f"{name}_{k}",
                        "source": f"source/framework/algorithm/test/{fname}",
                        "kind": op,
                        "element": element,
                        "array_shape": list(flow_direction.array_shape),
                        "partition_shape": list(flow_direction.partition_shape),
                        "ldd": to_json_raster(flow_direction),
                        "material": to_json_raster(material),
                        "expected": {"material": {"raster": to_json_raster(last[want_name]), "compare": "equal"}},
                    })
                elif isinstance(key, str):
                    last[key] = value
        out[op] = cases

    # downstream_distance: the case declares `DistanceElement const cell_size{N}` and
    # `q{std::sqrt(DistanceElement{2}) * cell_size}` with DistanceElement = float; q is evaluated here as the
    # test defines it (float32 arithmetic), the rasters come from the literals
    fname = "downstream_distance_test.cpp"
    cases = []
    for name, body in case_bodies(os.path.join(test_dir, fname)):
        stmts = split_statements(body)
        env = {}
        rest = []
        cell_size = None
        for stmt in stmts:
            vals = [t[1] for t in stmt]
            if "cell_size" in vals[:4] and vals[:2] == ["DistanceElement", "const"]:
                cell_size = float(vals[vals.index("{") + 1])
            elif vals[:3] == ["auto", "const", "q"]:
                continue
            else:
                rest.append(stmt)
        import numpy as np  # float32 arithmetic of the test's own expression

        env["q"] = float(np.sqrt(np.float32(2)) * np.float32(cell_size))
        run_statements(rest, env, [])
        flow_direction = env["flow_direction"]
        cases.append({
            "name": name,
            "source": f"source/framework/algorithm/test/{fname}",
            "kind": "downstream_distance",
            "element": "float32",
            "policies": "value",
            "cell_size": cell_size,
            "array_shape": list(flow_direction.array_shape),
            "partition_shape": list(flow_direction.partition_shape),
            "ldd": to_json_raster(flow_direction),
            "expected": {"distance": {"raster": to_json_raster(env["result_we_want"]), "compare": "equal"}},
        })
    out["downstream_distance"] = cases

    # inflow count: test_inflow_count([policies,] flow_direction, expected)
    path = os.path.join(test_dir, "inflow_count_test.cpp")
    text_toks = tokenize(strip_comments(open(path).read()))
    cases = []
    for name, body in case_bodies(path):
        env = {"array_shape": [3, 3], "partition_shape": [3, 3], "x": 255.0}
        env.update({k: float(v) for k, v in LDD_NAMES.items()})
        record = run_statements(split_statements(body), env, [])
        for k, v in record:
            if k == "test_inflow_count":
                rasters = [a for a in v if isinstance(a, Raster)]
                if len(rasters) < 2:
                    names = [a for a in v if isinstance(a, str)]
                    raise ValueError(name)
                ldd, want = rasters[-2], rasters[-1]
                cases.append({
                    "name": name,
                    "source": "source/framework/algorithm/test/inflow_count_test.cpp",
                    "kind": "inflow_count",
                    "array_shape": list(ldd.array_shape),
                    "partition_shape": list(ldd.partition_shape),
                    "ldd": to_json_raster(ldd),
                    "expected": {"count": {"raster": to_json_raster(want), "compare": "equal"}},
                })
    del text_toks
    out["inflow_count"] = cases

    # accu_info: the ldd fixture and the expected cell classes of each case
    fname = "accu_info_test.cpp"
    cases = []
    for name, body in case_bodies(os.path.join(test_dir, fname)):
        if name == "overloads":
            continue
        env = {"array_shape": [9, 9], "partition_shape": [3, 3]}
        record = run_statements(split_statements(body), env, [])
        checks = [v for k, v in record if k == "equal"]
        assert len(checks) == 1, name
        got_name, want_name = checks[0]
        _, opname, args = env[got_name]
        assert opname == "accu_info"
        ldd = args[0]
        cases.append({
            "name": name,
            "source": f"source/framework/algorithm/test/{fname}",
            "kind": "accu_info",
            "array_shape": list(ldd.array_shape),
            "partition_shape": list(ldd.partition_shape),
            "ldd": to_json_raster(ldd),
            "expected": {"cell_class": {"raster": to_json_raster(env[want_name]), "compare": "equal"}},
        })
    out["accu_info"] = cases
    return out


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    vectors = extract(ref_root)
    for op, cases in vectors.items():
        path = os.path.join(here, f"{op}.json")
        with open(path, "w") as f:
            json.dump({"op": op, "cases": cases}, f, indent=1)
        print(f"{op}: {len(cases)} cases -> {os.path.relpath(path)}")


if __name__ == "__main__":
    main()
