"""Extract the Joe-Kuo new-joe-kuo-6.21201 Sobol direction-number records.

The records are third-party data (Copyright (c) 2008 Frances Y. Kuo and
Stephen Joe, BSD-style licence, https://web.maths.unsw.edu.au/~fkuo/sobol/).
kiyosi bundles them as a flat ``u32`` array in
``crates/kiyosi/src/engines/monte_carlo/sobol_joe_kuo.rs``: each record is
``degree, coefficient bits, m_1 .. m_degree`` for Sobol dimensions 2..21201.

This script reads that array as text and writes it unchanged, as little-endian
``uint32`` words, to ``joe_kuo_6_21201.u32`` next to this file. It is run once,
by hand, where the kiyosi sources are available; the output file is what the
library and its tests read.

Usage: python extract_joe_kuo.py /path/to/kiyosi
"""

import pathlib
import re
import struct
import sys

RECORDS = 21200


def extract(root: pathlib.Path) -> list:
    text = (root / "crates/kiyosi/src/engines/monte_carlo/sobol_joe_kuo.rs").read_text()
    body = text[text.index("= [") + 3 : text.rindex("];")]
    words = [int(token) for token in re.findall(r"\d+", body)]
    # Walk the records to check their count and framing before writing.
    position, records = 0, 0
    while position < len(words):
        degree = words[position]
        position += degree + 2
        records += 1
    if position != len(words) or records != RECORDS:
        raise SystemExit(f"unexpected table layout: {records} records, {len(words)} words")
    return words


def main() -> None:
    root = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    words = extract(root)
    out = pathlib.Path(__file__).with_name("joe_kuo_6_21201.u32")
    out.write_bytes(struct.pack(f"<{len(words)}I", *words))
    print(f"wrote {len(words)} words ({RECORDS} records) to {out}")


if __name__ == "__main__":
    main()
