"""Regenerates rng_golden.json from libstdc++ by compiling and running gen_rng_golden.cpp."""
import json
import subprocess
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent

with tempfile.TemporaryDirectory() as d:
    exe = Path(d) / "gen"
    subprocess.run(["g++", "-O2", "-std=c++17", str(HERE / "gen_rng_golden.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
data = json.loads(out)
(HERE / "rng_golden.json").write_text(json.dumps(data, indent=0) + "\n")
print("wrote", HERE / "rng_golden.json")
