#!/usr/bin/env bash
# Regenerates tests/golden/ref_cpu/: outputs of the reference's own CPU
# programs (built from their sources by oracle/Makefile into oracle/_ref/),
# run on synthetic inputs placed at the paths they hard-code.
#   lensing  tables/lensing/std      ppm_synth lensing (u 0.3-0.5 x 256, 512)
#   response tables/ism_inst/std     ppm_synth response 270 128 seed 0
#   NSX      ext/model_data/nsx_H_v200804.out   ppm_synth nsx seed 1
# t_interp's 256 x 1024 flux matrices are kept at every 8th phase bin (and the
# last) and every 16th energy (and the last), as printed.
# Needs /root/reference (this container only).  Takes ~15 minutes.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REPO="$(cd "$HERE/../.." && pwd)"
RUN="$REPO/tests/_data/refrun"
SYNTH="$REPO/gpu_ppm_amd/_bin/ppm_synth"
make -s -C "$REPO/oracle" ref
mkdir -p "$RUN/output"
cd "$RUN"
"$SYNTH" lensing tables/lensing/std 0.3 0.5 256 512
"$SYNTH" response tables/ism_inst/std 270 128 0
"$SYNTH" nsx ext/model_data/nsx_H_v200804.out 1
"$REPO/oracle/_ref/os1_cpu" > os1.log
# the reference prints a warning per clamped table look-up; drop those lines
"$REPO/oracle/_ref/devarshi24st_cpu" | grep -v "out of bounds" > dev.log || true
"$REPO/oracle/_ref/t_interp" | grep -v "out of bounds" > t_interp.log || true
mkdir -p "$HERE/ref_cpu"
for f in output/os1?_cpu.txt; do cp "$f" "$HERE/ref_cpu/"; done
for f in output/*_counts_cpu.txt; do gzip -9 -c "$f" > "$HERE/ref_cpu/$(basename "$f").gz"; done
python3 - "$HERE/ref_cpu" <<'PY'
import gzip, sys
rows = list(range(0, 256, 8)) + [255]
cols = list(range(0, 1024, 16)) + [1023]
for t in ("10", "11", "12", "13", "20", "21", "22", "23"):
    lines = open("output/test%s_flux.txt" % t).read().split("\n")
    with gzip.open("%s/t_interp_test%s_sub.txt.gz" % (sys.argv[1], t), "wt") as f:
        f.write("# rows " + " ".join(map(str, rows)) + "\n# cols " + " ".join(map(str, cols)) + "\n")
        for r in rows:
            tok = lines[r].split()
            f.write(" ".join(tok[c] for c in cols) + "\n")
PY
sha256sum tables/lensing/std/* tables/ism_inst/std/* ext/model_data/nsx_H_v200804.out \
  | sed "s| tables/| tables/|" > "$HERE/ref_cpu/INPUTS.sha256"
