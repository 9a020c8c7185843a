"""Build the Graph<> facade drivers (host C++, g++) against librisgraph_amd.so.

Produces risgraph_amd/drivers/bin/{bfs,sssp,sswp,wcc}_inc and *_inc_batch,
the equivalents of the reference drivers of the same names.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(HERE, "bin")
LIBDIR = os.path.dirname(HERE)


def build():
    os.makedirs(BIN, exist_ok=True)
    outs = []
    for algo in ("bfs", "sssp", "sswp", "wcc"):
        for batch in (False, True):
            name = f"{algo}_inc" + ("_batch" if batch else "")
            out = os.path.join(BIN, name)
            cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-DALGO_{algo.upper()}",
                   *(["-DBATCH"] if batch else []), f"-I{os.path.join(ROOT, 'include')}",
                   os.path.join(HERE, "inc.cpp"), "-o", out, f"-L{LIBDIR}", "-lrisgraph_amd",
                   f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath,$ORIGIN/../.."]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stderr[-4000:]}")
            outs.append(out)
    return outs


if __name__ == "__main__":
    for o in build():
        print(o)
