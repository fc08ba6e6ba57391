"""Writes the source prelude of the per-program coefficient kernels compiled
at plan creation: include/slx.h, csrc/kernels/device.h and
csrc/kernels/coef_device.h inlined into one C++ raw string (local includes,
include guards and the C library headers, which the runtime compiler does not
ship, replaced by the fixed-width integer types).

usage: python3 gen_jit_prelude.py <output .inc>
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = [os.path.join(HERE, "..", "..", "include", "slx.h"),
         os.path.join(HERE, "kernels", "device.h"),
         os.path.join(HERE, "kernels", "coef_device.h")]
TYPES = """typedef signed char int8_t;
typedef unsigned char uint8_t;
typedef short int16_t;
typedef unsigned short uint16_t;
typedef int int32_t;
typedef unsigned int uint32_t;
typedef long int64_t;
typedef unsigned long uint64_t;
typedef __SIZE_TYPE__ size_t;
"""


def main():
    out = [TYPES]
    for f in FILES:
        for line in open(f):
            t = line.strip()
            if t.startswith("#include") or t == "#pragma once":
                continue
            out.append(line)
    text = "".join(out)
    assert ')SLXJIT"' not in text
    with open(sys.argv[1], "w") as fo:
        fo.write('static const char kJitPrelude[] = R"SLXJIT(' + text + ')SLXJIT";\n')


if __name__ == "__main__":
    main()
