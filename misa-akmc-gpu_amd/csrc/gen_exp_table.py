#!/usr/bin/env python3
"""Generate akmc_exp_table.h: the 2^(k/128) table of the table-driven exp.

The table layout follows the published exp design of ARM optimized-routines
(also glibc >= 2.28, sysdeps/ieee754/dbl-64/e_exp_data.c): for k in [0,128),
2^(k/128) ~= H[k] * (1 + T[k]) with H[k] the double nearest to 2^(k/128) and
T[k] the double nearest to the relative residual; the table stores
asuint64(T[k]) and asuint64(H[k]) - (k << 52)/128.  The values are derived
here from the mathematics (80-digit decimal arithmetic), not copied.
"""
import struct
import sys
from decimal import Decimal, getcontext

getcontext().prec = 80
N = 128


def main(path):
    ln2 = Decimal(2).ln()
    rows = []
    for k in range(N):
        exact = (ln2 * k / N).exp()
        h = float(exact)                      # correctly rounded Decimal -> double
        t = float((exact - Decimal(h)) / Decimal(h))
        hb = struct.unpack('<Q', struct.pack('<d', h))[0]
        tb = struct.unpack('<Q', struct.pack('<d', t))[0]
        rows.append((tb, (hb - ((k << 52) // N)) & 0xFFFFFFFFFFFFFFFF))
    with open(path, 'w') as f:
        f.write('// Generated by gen_exp_table.py -- do not edit.\n')
        f.write('#pragma once\n#include <stdint.h>\n')
        f.write('#define AKMC_EXP_TABLE_ENTRIES %d\n' % (2 * N))
        f.write('#define AKMC_EXP_TAB_INIT { \\\n')
        for tb, hb in rows:
            f.write('  0x%016xULL, 0x%016xULL, \\\n' % (tb, hb))
        f.write('}\n')
        f.write('static const uint64_t akmc_exp_tab_init[2 * 128] = AKMC_EXP_TAB_INIT;\n')


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else 'akmc_exp_table.h')
