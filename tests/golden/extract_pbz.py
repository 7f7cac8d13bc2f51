"""Write tests/golden/kpm_ref_pbz.npz from the reference's KPM regression fixtures
(tests/baseline_data/kpm/*.pbz: gzip-compressed pickles of one float32 numpy array each).

The pickles are never loaded: the opcode stream is read with `pickletools.genops`, which
executes nothing, and the array is rebuilt from its dtype string, shape and raw bytes.  Any
stream that is not exactly one plain numpy array is refused.

    python tests/golden/extract_pbz.py [/root/reference/tests/baseline_data/kpm]
"""
import glob
import gzip
import os
import pickletools
import sys

import numpy as np

ALLOWED_GLOBALS = {("numpy.core.multiarray", "_reconstruct"), ("numpy", "ndarray"), ("numpy", "dtype")}


def parse_array(data):
    ops = list(pickletools.genops(data))
    strings, memo, last = [], [], None
    for op, arg, _ in ops:
        if op.name == "STACK_GLOBAL":
            pair = tuple(strings[-2:])
            if pair not in ALLOWED_GLOBALS:
                raise ValueError("unexpected global {}".format(pair))
        if op.name in ("GLOBAL", "INST", "OBJ", "NEWOBJ", "NEWOBJ_EX", "EXT1", "EXT2", "EXT4", "PERSID", "BINPERSID"):
            raise ValueError("unexpected opcode {}".format(op.name))
        if op.name in ("SHORT_BINUNICODE", "BINUNICODE", "UNICODE"):
            strings.append(arg)
            last = arg
        elif op.name == "MEMOIZE":
            memo.append(last)
        elif op.name in ("BINGET", "LONG_BINGET"):
            last = memo[arg]
            if isinstance(last, str):
                strings.append(last)
        else:
            last = None
    names = [a for o, a, _ in ops if o.name in ("SHORT_BINUNICODE", "BINUNICODE")]
    dtype = names[names.index("dtype") + 1]
    order = [a for o, a, _ in ops if o.name in ("SHORT_BINUNICODE",) and a in ("<", ">", "|", "=")][0]
    raw = [a for o, a, _ in ops if o.name in ("SHORT_BINBYTES", "BINBYTES", "BINBYTES8")][-1]
    fortran = [o.name for o, a, _ in ops if o.name in ("NEWFALSE", "NEWTRUE")][-1] == "NEWTRUE"
    # shape: the ints after the state's version number (1), up to the first TUPLE opcode
    names_ops = [(o.name, a) for o, a, _ in ops]
    start = next(i for i, (n, a) in enumerate(names_ops)
                 if n == "MARK" and i > 0 and any(m == "REDUCE" for m, _ in names_ops[:i]))
    shape = []
    i = start + 2   # skip MARK and the version
    while names_ops[i][0] in ("BININT1", "BININT2", "BININT", "LONG1"):
        shape.append(int(names_ops[i][1]))
        i += 1
    if names_ops[i][0] not in ("TUPLE1", "TUPLE2", "TUPLE3", "TUPLE", "EMPTY_TUPLE"):
        raise ValueError("unexpected shape encoding")
    arr = np.frombuffer(raw, dtype=np.dtype(order + dtype))
    return arr.reshape(shape, order="F" if fortran else "C").copy()


def main(src):
    out = {}
    for path in sorted(glob.glob(os.path.join(src, "*.pbz"))):
        name = os.path.basename(path)[:-4]
        with gzip.open(path, "rb") as f:
            out[name] = parse_array(f.read())
        print(name, out[name].dtype, out[name].shape, out[name].ravel()[:3])
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kpm_ref_pbz.npz")
    np.savez(dst, **out)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tests/baseline_data/kpm")
