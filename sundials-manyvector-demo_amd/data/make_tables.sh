#!/bin/sh
# Regenerate primordial_tables.f64 and primordial_tables.json from the reference's
# HDF5 table file. Needs the HDF5 C library (found under /opt/conda in the build
# container); the outputs are committed, so nothing else ever runs this.
set -e
here=$(cd "$(dirname "$0")" && pwd)
h5=${1:-/root/reference/tests/primordial_ode/primordial_tables.h5}
tmp=$(mktemp -d)
gcc -O2 -o "$tmp/extract_tables" "$here/extract_tables.c" \
    -I/opt/conda/include -L/opt/conda/lib -lhdf5_hl -lhdf5 -Wl,-rpath,/opt/conda/lib
"$tmp/extract_tables" "$h5" "$here/primordial_tables.f64" > "$tmp/names.txt"
python3 - "$here" "$tmp/names.txt" "$h5" <<'EOF'
import hashlib, json, sys, os
here, names_file, h5 = sys.argv[1:4]
names = [l.strip() for l in open(names_file) if l.strip()]
blob = open(os.path.join(here, "primordial_tables.f64"), "rb").read()
assert len(blob) == len(names) * 1024 * 8
man = {
    "source": "tests/primordial_ode/primordial_tables.h5",
    "source_md5": hashlib.md5(open(h5, "rb").read()).hexdigest(),
    "dtype": "float64-le", "table_len": 1024, "order": names,
    "sha256": hashlib.sha256(blob).hexdigest(),
    "table_sha256": {n: hashlib.sha256(blob[i*8192:(i+1)*8192]).hexdigest()
                     for i, n in enumerate(names)},
}
json.dump(man, open(os.path.join(here, "primordial_tables.json"), "w"), indent=1)
print("wrote", len(names), "tables")
EOF
rm -rf "$tmp"
