#!/usr/bin/env python3
"""Extract the cluster-fit ordering tables of the reference BC1 encoder as data.

The BC1 cluster-fit stage walks a list of "likely best" selector orderings per
starting ordering.  Those lists were produced offline by training
(rgbcx), so they cannot be derived from a formula: they are carried here as
data, extracted verbatim (values only) from the reference's table sources:

  * ``OrderTable<3>::Orders``      153 x 3 histograms   quicktex/s3tc/bc1/OrderTable.cpp:46-61
  * ``OrderTable<4>::Orders``      969 x 4 histograms   quicktex/s3tc/bc1/OrderTable.cpp:63-163
  * ``OrderTable<3>::BestOrders``  153 x 32 u16          quicktex/s3tc/bc1/OrderTable.cpp:165-321
  * ``OrderTable<4>::BestOrders``  969 x 128 u16         quicktex/s3tc/bc1/OrderTable4.cpp:996-1966
    (the 128-wide branch; ``RGBCX_USE_SMALLER_TABLES`` is undefined in the reference build)
  * ``OrderTable<3/4>::SingleColorHashes``: the ordering a histogram with all 16
    selectors in bin i starts from (quicktex/s3tc/bc1/OrderTable.cpp:43-44).  Data,
    not derived: the 3-colour list is {12, 15, 89} although ordering 15 is the
    histogram (0, 0, 16) and 89 is (0, 16, 0), i.e. bins 1 and 2 map crosswise,
    and the encoder's output depends on it.

Everything else the encoder needs from ``OrderTable`` (hash lookup, least
squares factors) is recomputed from these histograms by the library itself.

Run once in a container that holds the reference sources::

    python3 quicktex_amd/csrc/gen_order_tables.py /root/reference/quicktex/s3tc/bc1

It writes ``order_tables.inc`` next to this script.  The generated file records
a SHA-256 over the table values; ``tests/test_tables.py`` re-derives the
structural invariants (every histogram sums to 16, all 969/153 compositions
present exactly once, every best-order entry in range).
"""
import hashlib
import os
import re
import sys


def _region(text, start_marker, end_marker):
    i = text.index(start_marker)
    j = text.index(end_marker, i)
    return text[i:j]


def _ints(chunk):
    # strip comments, then take every integer literal
    chunk = re.sub(r'//[^\n]*', '', chunk)
    return [int(x) for x in re.findall(r'\b\d+\b', chunk)]


def extract(bc1_dir):
    ot = open(os.path.join(bc1_dir, 'OrderTable.cpp')).read()
    ot4 = open(os.path.join(bc1_dir, 'OrderTable4.cpp')).read()

    orders3 = _ints(_region(ot, 'OrderTable<3>::Orders = {', '// endregion').split('= {', 1)[1])
    orders4 = _ints(_region(ot, 'OrderTable<4>::Orders = {', '// endregion').split('= {', 1)[1])
    best3 = _ints(_region(ot, 'OrderTable<3>::BestOrders = {', '// endregion').split('= {', 1)[1])
    big = ot4.split('OrderTable<4>::BestOrders = {{', 1)[1]
    big = big.split('#else', 1)[1].split('#endif', 1)[0]
    best4 = _ints(big)

    m3 = re.search(r'OrderTable<3>::SingleColorHashes = \{([^}]*)\}', ot)
    m4 = re.search(r'OrderTable<4>::SingleColorHashes = \{([^}]*)\}', ot)
    single3 = [int(x) for x in m3.group(1).split(',')]
    single4 = [int(x) for x in m4.group(1).split(',')]
    assert len(single3) == 3 and len(single4) == 4
    assert len(orders3) == 153 * 3, len(orders3)
    assert len(orders4) == 969 * 4, len(orders4)
    assert len(best3) == 153 * 32, len(best3)
    assert len(best4) == 969 * 128, len(best4)
    return orders3, orders4, best3, best4, single3, single4


def digest(orders3, orders4, best3, best4):
    h = hashlib.sha256()
    h.update(bytes(orders3))
    h.update(bytes(orders4))
    for v in best3 + best4:
        h.update(v.to_bytes(2, 'little'))
    return h.hexdigest()


def _rows(vals, width, per_line):
    out = []
    for i in range(0, len(vals), per_line):
        out.append('  ' + ','.join(str(v) for v in vals[i:i + per_line]) + ',')
    return '\n'.join(out)


def main():
    bc1_dir = sys.argv[1] if len(sys.argv) > 1 else '/root/reference/quicktex/s3tc/bc1'
    orders3, orders4, best3, best4, single3, single4 = extract(bc1_dir)
    sha = digest(orders3, orders4, best3, best4)
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, 'order_tables.inc'), 'w') as f:
        f.write('// Generated by gen_order_tables.py -- cluster-fit ordering data (values only).\n')
        f.write('// Source: quicktex/s3tc/bc1/OrderTable.cpp:46-321, OrderTable4.cpp:996-1966 (Apache-2.0).\n')
        f.write('// SHA-256 over (orders3 u8, orders4 u8, best3 u16le, best4 u16le): %s\n' % sha)
        f.write('#define QTX_ORDER_TABLES_SHA256 "%s"\n' % sha)
        f.write('static const unsigned char qtx_orders3[153 * 3] = {\n%s\n};\n' % _rows(orders3, 3, 48))
        f.write('static const unsigned char qtx_orders4[969 * 4] = {\n%s\n};\n' % _rows(orders4, 4, 48))
        f.write('static const unsigned short qtx_best_orders3[153 * 32] = {\n%s\n};\n' % _rows(best3, 32, 32))
        f.write('static const unsigned short qtx_best_orders4[969 * 128] = {\n%s\n};\n' % _rows(best4, 128, 32))
        f.write('// ordering index for a histogram with all 16 selectors in bin i (3-colour: bins 1, 2 crosswise)\n')
        f.write('static const unsigned short qtx_single_color_hashes3[3] = {%s};\n' % ', '.join(map(str, single3)))
        f.write('static const unsigned short qtx_single_color_hashes4[4] = {%s};\n' % ', '.join(map(str, single4)))
    print('wrote order_tables.inc sha256', sha)


if __name__ == '__main__':
    main()
