"""Mutation check of the oracle's simulate pass (test infrastructure).

Plants one mistake at a time in a copy of firex_oracle.cpp's simulate
(shaders.ts:151-342 restated): a changed constant, a flipped sign, a swapped
operand or a dropped term, builds that copy into a temporary directory and
runs the CPU pinning tests against it (FIREX_ORACLE_LIB).  A mutant is
"killed" when at least one test fails.  The working tree is not touched.

    python oracle/mutants.py [-j JOBS] [--only NAME ...] [--tests FILE ...]

Prints one line per mutant and a JSON summary at the end; exits 1 if any
mutant survives.
"""
from __future__ import annotations

import argparse
import concurrent.futures
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCE = os.path.join(HERE, "firex_oracle.cpp")
CXXFLAGS = ["-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-mfma", "-shared"]
TESTS = ["tests/test_oracle_simulate_terms.py", "tests/test_oracle_pinning.py"]

# (name, text in simulate, replacement); each text must occur exactly once in
# the simulate function.
MUTANTS = [
    # six-neighbour blends, shaders.ts:177-187
    ("diffusion velocity 0.32->0.23", "dt * T(0.32)", "dt * T(0.23)"),
    ("diffusion velocity cap 0.12->0.13", "T(0.32), T(0), T(0.12)", "T(0.32), T(0), T(0.13)"),
    ("diffusion temperature 0.18->0.17", "dt * T(0.18)", "dt * T(0.17)"),
    ("diffusion temperature cap 0.08->0.09", "T(0.18), T(0), T(0.08)", "T(0.18), T(0), T(0.09)"),
    ("diffusion species 0.12->0.13", "dt * T(0.12), T(0), T(0.06)", "dt * T(0.13), T(0), T(0.06)"),
    ("diffusion species cap 0.06->0.07", "dt * T(0.12), T(0), T(0.06)", "dt * T(0.12), T(0), T(0.07)"),
    ("velocity mean missing a neighbour", "L.x + R.x + B.x + Tp.x + K.x + F.x", "L.x + R.x + B.x + Tp.x + K.x + K.x"),
    ("temperature mean /6->/5", "L.w + R.w + B.w + Tp.w + K.w + F.w, T(6)", "L.w + R.w + B.w + Tp.w + K.w + F.w, T(5)"),
    ("species mean uses a flow neighbour", "sL.z + sR.z + sB.z + sT.z + sK.z + sF.z",
     "sL.z + sR.z + sB.z + sT.z + sK.z + sK.z"),
    # source injection, swirl and lift, shaders.ts:189-201
    ("thermal 0.22->0.2", "T(1) + T(0.22) * (powerRoot - T(1))", "T(1) + T(0.2) * (powerRoot - T(1))"),
    ("heat injection 1180->1100", "source * T(1180)", "source * T(1100)"),
    ("fuel injection uses powerRoot", "source * chemX * firePower", "source * chemX * powerRoot"),
    ("swirl 0.11->0.13", "dt * T(0.11)", "dt * T(0.13)"),
    ("swirl phase 2.3->2.2", "time * T(2.3)", "time * T(2.2)"),
    ("swirl frequency 22->21", "npz * T(22)", "npz * T(21)"),
    ("swirl phase 1.9->2.0", "time * T(1.9)", "time * T(2.0)"),
    ("swirl frequency 25->24", "npx * T(25)", "npx * T(24)"),
    ("lift 0.52->0.5", "dt * T(0.52)", "dt * T(0.5)"),
    # ambient walls, shaders.ts:203-208
    ("wall relaxation 0.035->0.053", "ambient, T(0.035)", "ambient, T(0.053)"),
    ("floor counted as ambient wall", "bool ambientBoundary = x == 0 ||", "bool ambientBoundary = y == 0 || x == 0 ||"),
    ("wall oxygen not restored", "sp.y = wmax(sp.y, chemY);", "sp.y = sp.y;"),
    # liquid, shaders.ts:210-215, 290-297
    ("water scale 3.4->3.3", "liq * T(3.4)", "liq * T(3.3)"),
    ("water cap 5->4", "T(3.4), T(0), T(5)", "T(3.4), T(0), T(4)"),
    ("evaporation edge 335->345", "wsmooth(T(335), T(1040)", "wsmooth(T(345), T(1040)"),
    ("evaporation edge 1040->1000", "wsmooth(T(335), T(1040)", "wsmooth(T(335), T(1000)"),
    ("evaporation rate 13.5->12.5", "dt * T(13.5)", "dt * T(12.5)"),
    ("evaporation heat 2080->2800", "evaporation * T(2080)", "evaporation * T(2800)"),
    ("liquid cooling 260->250", "dt * T(260)", "dt * T(250)"),
    ("evaporated vapor 2.15->2.5", "evaporation * T(2.15)", "evaporation * T(2.5)"),
    ("water product decay 4.8->4.2", "dt * T(4.8)", "dt * T(4.2)"),
    ("spray momentum 0.82->0.8", "smx * water * dt * T(0.82)", "smx * water * dt * T(0.8)"),
    ("spray nozzle x 0.035->0.05", "T(0.035) * ex", "T(0.05) * ex"),
    # entrainment, shaders.ts:217-232
    ("entrainment speed 0.28->0.82", "T(0.28) * speed", "T(0.82) * speed"),
    ("entrainment base 0.12->0.21", "(T(0.12) + T(0.28)", "(T(0.21) + T(0.28)"),
    ("entrainment heat 0.22->0.2", "T(0.22) * wsmooth(T(400)", "T(0.2) * wsmooth(T(400)"),
    ("entrainment heat edge 400->450", "wsmooth(T(400), T(1500)", "wsmooth(T(450), T(1500)"),
    ("fuel core 0.30->0.35", "wsmooth(T(0.30), T(0.78)", "wsmooth(T(0.35), T(0.78)"),
    ("fuel core 0.78->0.8", "wsmooth(T(0.30), T(0.78)", "wsmooth(T(0.30), T(0.8)"),
    ("interface weight 0.18->0.28", "wmix(T(1), T(0.18), fuelCore)", "wmix(T(1), T(0.28), fuelCore)"),
    ("entrainment cap 0.08->0.1", "interfaceWeight,\n              T(0), T(0.08))",
     "interfaceWeight,\n              T(0), T(0.1))"),
    ("deficit not normalised", "wdivr(ambientOxygen - sp.y, ambientOxygen)", "(ambientOxygen - sp.y)"),
    # reaction, shaders.ts:234-256
    ("ignition edge 690->700", "wsmooth(T(690), T(930)", "wsmooth(T(700), T(930)"),
    ("ignition edge 930->900", "wsmooth(T(690), T(930)", "wsmooth(T(690), T(900)"),
    ("activation 15000->16000", "-T(15000)", "-T(16000)"),
    ("Arrhenius 2.5e6->2.4e6", "T(2.5e6)", "T(2.4e6)"),
    ("fuel order -0.30->-0.25", "T(-0.30)", "T(-0.25)"),
    ("oxygen order 1.30->1.25", "T(1.30)", "T(1.25)"),
    ("rate cap 80->60", "T(0), T(80)", "T(0), T(60)"),
    ("available fuel 0.25->0.2", "oxygenBefore * T(0.25)", "oxygenBefore * T(0.2)"),
    ("burn water 18->16", "-water * T(18)", "-water * T(16)"),
    ("oxygen per fuel 4->3.5", "oxygenBefore - burn * T(4)", "oxygenBefore - burn * T(3.5)"),
    ("richness 1.8->1.6", "richness * T(1.8)", "richness * T(1.6)"),
    ("burn vapor 0.62->0.6", "burn * T(0.62)", "burn * T(0.6)"),
    ("heat release 2150->2100", "T(2150) + T(1450)", "T(2100) + T(1450)"),
    ("heat release 1450->1400", "T(2150) + T(1450)", "T(2150) + T(1400)"),
    ("reaction diagnostic 0.22->0.2", "wmax(dt, T(0.0001))) * T(0.22)", "wmax(dt, T(0.0001))) * T(0.2)"),
    # vorticity-like force, shaders.ts:258-272
    ("cross product reversed", "(diry * curlz - dirz * curly) * vk", "(dirz * curly - diry * curlz) * vk"),
    ("cross y reversed", "(dirz * curlx - dirx * curlz) * vk", "(dirx * curlz - dirz * curlx) * vk"),
    ("cross z reversed", "(dirx * curly - diry * curlx) * vk", "(diry * curlx - dirx * curly) * vk"),
    ("vorticity 0.00135->0.0015", "T(0.00135)", "T(0.0015)"),
    ("curl x sign", "(Tp.z - B.z) * my - (F.y - K.y) * mz", "(Tp.z - B.z) * my + (F.y - K.y) * mz"),
    ("curl y metric", "(F.x - K.x) * mz - (R.z - L.z) * mx", "(F.x - K.x) * mx - (R.z - L.z) * mx"),
    ("curl z operands", "(R.y - L.y) * mx - (Tp.x - B.x) * my", "(R.y - L.y) * mx - (Tp.z - B.z) * my"),
    ("gradient metric", "T(0.5) * ((R.w - L.w) * mx)", "T(0.5) * ((R.w - L.w) * my)"),
    # turbulence, shaders.ts:274-285
    ("turbulence 0.16->0.61", "T(0.075) + T(0.16) * chemW", "T(0.075) + T(0.61) * chemW"),
    ("turbulence 0.075->0.07", "T(0.075) + T(0.16) * chemW", "T(0.07) + T(0.16) * chemW"),
    ("hot plume 520->500", "wsmooth(T(520), T(1650)", "wsmooth(T(500), T(1650)"),
    ("hot plume 1650->1600", "wsmooth(T(520), T(1650)", "wsmooth(T(520), T(1600)"),
    ("envelope 0.045->0.05", "wsmooth(T(0.045), T(0.15)", "wsmooth(T(0.05), T(0.15)"),
    ("envelope 0.80->0.85", "wsmooth(T(0.80), T(1.0)", "wsmooth(T(0.85), T(1.0)"),
    ("ltx frequency 31->30", "npy * T(31)", "npy * T(30)"),
    ("ltx phase 3.1->3.0", "time * T(3.1)", "time * T(3.0)"),
    ("ltx second amplitude 0.45->0.4", "T(0.45) * Math<T>::sin", "T(0.4) * Math<T>::sin"),
    ("ltx second phase 5.3->5.0", "time * T(5.3)", "time * T(5.0)"),
    ("lty amplitude 0.12->0.13", "T(0.12) * Math<T>::sin(npx", "T(0.13) * Math<T>::sin(npx"),
    ("lty phase 4.1->4.0", "time * T(4.1)", "time * T(4.0)"),
    ("ltz frequency 21->20", "npx * T(21)", "npx * T(20)"),
    ("ltz phase 2.7->2.5", "time * T(2.7)", "time * T(2.5)"),
    ("ltz second frequency 59->58", "npy * T(59)", "npy * T(58)"),
    ("ltz second phase 4.7->4.5", "time * T(4.7)", "time * T(4.5)"),
    ("turbulence y uses ltx", "fl.y = fl.y + lty * tk;", "fl.y = fl.y + ltx * tk;"),
    # buoyancy, radiation, decays, shaders.ts:287-304
    ("buoyancy 1500->1400", "wdivr(fl.w - ambient, T(1500))", "wdivr(fl.w - ambient, T(1400))"),
    ("buoyancy 0.72->0.7", "dt * T(0.72)", "dt * T(0.7)"),
    ("buoyancy cap 1.2->1.1", "T(-0.1), T(1.2)", "T(-0.1), T(1.1)"),
    ("radiation 5.67e-8->5.0e-8", "T(5.67e-8)", "T(5.0e-8)"),
    ("radiation 1005->1000", "T(1005) * persistence", "T(1000) * persistence"),
    ("product decay 0.075->0.08", "-dt * T(0.075) / persistence", "-dt * T(0.08) / persistence"),
    ("vapor decay 0.14->0.15", "-dt * T(0.14) / persistence", "-dt * T(0.15) / persistence"),
    # outflow and top row, shaders.ts:306-321
    ("outflow 0.16->0.61", "outflow * T(0.16) / persistence", "outflow * T(0.61) / persistence"),
    ("outflow edge 0.84->0.85", "wsmooth(T(0.84), T(0.995)", "wsmooth(T(0.85), T(0.995)"),
    ("outflow edge 0.995->0.99", "wsmooth(T(0.84), T(0.995)", "wsmooth(T(0.84), T(0.99)"),
    ("outflow decay 15->12", "-outflow * dt * T(15)", "-outflow * dt * T(12)"),
    ("reaction not scaled by outflow", "react *= T(1) - outflow;", "react *= T(1);"),
    ("fuel not scaled by outflow", "sp.x *= T(1) - outflow;", "sp.x *= T(1);"),
    ("top row keeps its temperature", "fl.w = ambient;\n            react = T(0);", "react = T(0);"),
    # boundary conditions and clamps, shaders.ts:323-339
    ("velocity clamp 1.4->1.5", "fl.x = wclamp(fl.x, T(-1.4), T(1.4));", "fl.x = wclamp(fl.x, T(-1.5), T(1.5));"),
    ("temperature cap 2900->2800", "wclamp(fl.w, ambient, T(2900))", "wclamp(fl.w, ambient, T(2800))"),
    ("open top admits inflow", "if (y == ny - 1) fl.y = wmax(T(0), fl.y);", ""),
    ("floor keeps normal velocity", "if (y == 0) fl.y = T(0);", ""),
]


def simulate_span(text: str) -> tuple[int, int]:
    start = text.index("  void simulate(int sIn) {")
    end = text.index("  // ------------------------------------------------------------- droplets", start)
    return start, end


def mutate(text: str, old: str, new: str) -> str:
    a, b = simulate_span(text)
    body = text[a:b]
    n = body.count(old)
    if n != 1:
        raise ValueError(f"{old!r} occurs {n} times in simulate")
    return text[:a] + body.replace(old, new) + text[b:]


def run_one(mut, text: str, tests: list[str], workdir: str) -> dict:
    name, old, new = mut
    d = tempfile.mkdtemp(dir=workdir)
    src = os.path.join(d, "firex_oracle.cpp")
    lib = os.path.join(d, "liboracle_firex.so")
    with open(src, "w") as f:
        f.write(mutate(text, old, new))
    res = subprocess.run(["g++", *CXXFLAGS, "-I", HERE, "-o", lib, src], capture_output=True, text=True)
    if res.returncode != 0:
        return {"name": name, "status": "build-error", "detail": res.stderr[-400:]}
    env = dict(os.environ, FIREX_ORACLE_LIB=lib, OMP_NUM_THREADS="1")
    res = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *tests],
                         cwd=ROOT, env=env, capture_output=True, text=True)
    failed = [ln.split(" ", 1)[1].split(" - ")[0] for ln in res.stdout.splitlines() if ln.startswith("FAILED ")]
    return {"name": name, "status": "killed" if res.returncode != 0 else "survived",
            "by": failed[:1]}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", "--jobs", type=int, default=4)
    ap.add_argument("--only", nargs="*", default=None, help="mutant names to run (substring match)")
    ap.add_argument("--tests", nargs="*", default=TESTS)
    args = ap.parse_args()
    text = open(SOURCE).read()
    muts = [m for m in MUTANTS if not args.only or any(s in m[0] for s in args.only)]
    for m in muts:  # every mutant must apply before anything runs
        mutate(text, m[1], m[2])
    results = []
    with tempfile.TemporaryDirectory(prefix="firex_mutants_") as work:
        with concurrent.futures.ThreadPoolExecutor(max_workers=args.jobs) as pool:
            for r in pool.map(lambda m: run_one(m, text, args.tests, work), muts):
                print(f"{r['status']:11s} {r['name']:45s} {' '.join(r.get('by', []))}", flush=True)
                results.append(r)
    survived = [r["name"] for r in results if r["status"] != "killed"]
    print(json.dumps({"mutants": len(results), "killed": len(results) - len(survived), "survived": survived}))
    return 1 if survived else 0


if __name__ == "__main__":
    sys.exit(main())
