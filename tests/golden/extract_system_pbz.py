"""Write tests/golden/system_ref_pbz.npz from the reference's system regression fixtures
(tests/baseline_data/system/expected[...].pbz: gzip-compressed pickles of built `System` objects).

Nothing is unpickled and nothing is imported or called: the opcode stream from
`pickletools.genops` drives a small symbolic stack machine in which every global is an opaque
(module, name) pair, REDUCE builds an opaque call record, and BUILD attaches its state.  numpy
arrays are rebuilt from the records of `numpy.core.multiarray._reconstruct` (dtype string, shape,
raw bytes).  Opcodes outside that subset stop the extraction.

Kept per fixture: site positions (x, y, z), the sublattice id of every site, the hopping pairs
of every hopping family, and for periodic systems every boundary's shift and hopping pairs.

    python tests/golden/extract_system_pbz.py [/root/reference/tests/baseline_data/system]
"""
import glob
import gzip
import os
import pickletools
import sys

import numpy as np

KEEP = ["graphene-monolayer-periodic-1d", "graphene-monolayer-periodic-2d", "graphene-monolayer-nn",
        "graphene-monolayer-4atom", "graphene-monolayer-4atom-periodic-2d", "graphene-monolayer-periodic-1d-alt"]


class Global:
    def __init__(self, module, name):
        self.key = (module, name)


class Call:
    """REDUCE of a global with arguments; BUILD stores the state."""

    def __init__(self, fn, args):
        self.fn, self.args, self.state = fn, args, None


class Mark:
    pass


def run(data):
    stack, memo = [], {}

    def pop_mark():
        i = len(stack) - 1
        while not isinstance(stack[i], Mark):
            i -= 1
        items = stack[i + 1:]
        del stack[i:]
        return items

    for op, arg, _ in pickletools.genops(data):
        n = op.name
        if n in ("PROTO", "FRAME"):
            continue
        elif n == "STOP":
            break
        elif n in ("SHORT_BINUNICODE", "BINUNICODE", "UNICODE", "SHORT_BINBYTES", "BINBYTES", "BINBYTES8",
                   "BININT1", "BININT2", "BININT", "LONG1", "BINFLOAT"):
            stack.append(arg)
        elif n in ("NEWTRUE", "NEWFALSE"):
            stack.append(n == "NEWTRUE")
        elif n == "NONE":
            stack.append(None)
        elif n == "MEMOIZE":
            memo[len(memo)] = stack[-1]
        elif n in ("BINPUT", "LONG_BINPUT"):
            memo[arg] = stack[-1]
        elif n in ("BINGET", "LONG_BINGET"):
            stack.append(memo[arg])
        elif n == "MARK":
            stack.append(Mark())
        elif n == "STACK_GLOBAL":
            name = stack.pop()
            module = stack.pop()
            stack.append(Global(module, name))
        elif n == "EMPTY_TUPLE":
            stack.append(())
        elif n in ("TUPLE1", "TUPLE2", "TUPLE3"):
            k = int(n[-1])
            items = tuple(stack[-k:])
            del stack[-k:]
            stack.append(items)
        elif n == "TUPLE":
            stack.append(tuple(pop_mark()))
        elif n == "EMPTY_LIST":
            stack.append([])
        elif n == "EMPTY_DICT":
            stack.append({})
        elif n == "APPENDS":
            items = pop_mark()
            stack[-1].extend(items)
        elif n == "APPEND":
            v = stack.pop()
            stack[-1].append(v)
        elif n == "SETITEMS":
            items = pop_mark()
            for k in range(0, len(items), 2):
                stack[-1][items[k]] = items[k + 1]
        elif n == "SETITEM":
            v = stack.pop()
            k = stack.pop()
            stack[-1][k] = v
        elif n == "REDUCE":
            args = stack.pop()
            fn = stack.pop()
            stack.append(Call(fn, args))
        elif n == "NEWOBJ":
            args = stack.pop()
            cls = stack.pop()
            stack.append(Call(cls, args))
        elif n == "BUILD":
            state = stack.pop()
            stack[-1].state = state
        else:
            raise ValueError("opcode outside the extraction subset: " + n)
    return stack[-1]


def to_array(obj):
    """numpy array from a `_reconstruct` record whose state is (version, shape, dtype, fortran, raw)."""
    _, shape, dtype, fortran, raw = obj.state
    code = dtype.args[0]
    order = dtype.state[1]
    arr = np.frombuffer(raw, dtype=np.dtype((order if order in "<>" else "") + code))
    return arr.reshape(shape, order="F" if fortran else "C").copy()


def extract(system):
    """The arrays kept for one System record (two layouts: the C++ system as the whole state, or
    under the key `impl` of a dict state)."""
    impl = system.state["impl"] if isinstance(system.state, dict) else system.state
    st = impl.state
    out = {}
    for k, a in zip("xyz", st["positions"].state):
        out[k] = to_array(a)
    cs = st["compressed_sublattices"].state
    alias = to_array(cs["alias_ids"])
    counts = to_array(cs["site_counts"])
    out["sublattice"] = np.repeat(alias, counts).astype(np.int32)

    def blocks(hb, prefix):
        name_map = hb.state["name_map"]
        names = sorted(name_map, key=lambda k: name_map[k])
        for fam, (rows, cols) in zip(names, hb.state["data"]):
            out[prefix + fam + "_rows"] = to_array(rows).astype(np.int32)
            out[prefix + fam + "_cols"] = to_array(cols).astype(np.int32)

    blocks(st["hopping_blocks"], "hop_")
    for b, boundary in enumerate(st["boundaries"]):
        hb, shift = boundary.state
        blocks(hb, "b%d_" % b)
        out["b%d_shift" % b] = to_array(shift)
    return out


def main(src):
    out = {}
    for path in sorted(glob.glob(os.path.join(src, "*.pbz"))):
        name = os.path.basename(path)[len("expected["):-len("].pbz")]
        if name not in KEEP:
            continue
        with gzip.open(path, "rb") as f:
            arrays = extract(run(f.read()))
        for k, v in arrays.items():
            out[name + "/" + k] = v
        print(name, {k: v.shape for k, v in arrays.items()})
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "system_ref_pbz.npz")
    np.savez(dst, **out)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tests/baseline_data/system")
