"""Mutation check of the oracle's pins (CPU only; run by hand: python tests/mutation_check.py).

Each entry below is a plausible slip in the oracle's restatement of the reference shaders.  The
script applies one slip at a time to a scratch copy of the repository's oracle, rebuilds it, and
runs every tests/test_oracle_*.py file against it.  A slip that no test catches ("SURVIVED") marks
a part of the oracle that no CPU check pins.  Nothing here touches the repository itself."""
import os
import shutil
import subprocess
import sys
import tempfile

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# (name, text in oracle/ddgi_oracle.cpp (or the file named fourth), replacement) -- the shader lines each follows:
# sampleProbes src/shaders/irradiance.glsl:181-234, closest hit src/shaders/closesthit.glsl:137-285
MUTATIONS = [
    ("Chebyshev factor squared", "cheb = std::max(cheb * cheb * cheb, T(0));", "cheb = std::max(cheb * cheb, T(0));"),
    ("trilinear weights swapped",
     "const V3<T> tri(off[0] ? alpha.x : T(1) - alpha.x, off[1] ? alpha.y : T(1) - alpha.y,\n\t\t\t\t\t\toff[2] ? alpha.z : T(1) - alpha.z);",
     "const V3<T> tri(off[0] ? T(1) - alpha.x : alpha.x, off[1] ? T(1) - alpha.y : alpha.y,\n\t\t\t\t\t\toff[2] ? T(1) - alpha.z : alpha.z);"),
    ("back-face weight not squared", "weight *= bfw * bfw + T(0.2);", "weight *= bfw + T(0.2);"),
    ("crush threshold 0.1", "const T crush = T(0.2);", "const T crush = T(0.1);"),
    ("crush cubed -> squared", "if(weight < crush) weight *= weight * weight * (T(1) / (crush * crush));",
     "if(weight < crush) weight *= weight * (T(1) / crush);"),
    ("trilinear floor 0.001 dropped", "const T tw = tri.x * tri.y * tri.z + T(0.001);", "const T tw = tri.x * tri.y * tri.z;"),
    ("fallback mix x4 instead of x8", "T(8) * clampT(ctx.total_w, T(0), T(1) / T(8))", "T(4) * clampT(ctx.total_w, T(0), T(1) / T(4))"),
    ("back-face depth x1 instead of x0.8", "out[3] = t * T(0.8);", "out[3] = t;"),
    ("diffuse without (1 - f0)", "const V3<T> diffuse = albedo * (V3<T>(1, 1, 1) - f0) * (T(1) - metal);",
     "const V3<T> diffuse = albedo * (T(1) - metal);"),
    ("specular looked up along the normal",
     "const V3<T> reflection = sample_probes<T>(g, states, irr, dep, position, refl_dir, -d);",
     "const V3<T> reflection = sample_probes<T>(g, states, irr, dep, position, normal, -d);"),
    ("sun fade on the direct term only", "\t\tcolor = color * fade;\n\t}\n\tif(aux) {", "\t}\n\tif(aux) {"),
    # blend and state machine, src/shaders/probesUpdate.glsl:49-119
    ("hysteresis applied to the new value", "px[c] = double(mixT(acc[c], prev, h));", "px[c] = double(mixT(prev, acc[c], h));"),
    ("depth lobe without its exponent", "const T w = std::pow(std::max(T(0), dot(td, rd)), sharp);",
     "const T w = std::max(T(0), dot(td, rd));"),
    ("depth of a miss 0 instead of the cell diagonal", "\t\t\t\t\tif(depth < T(0)) depth = cell_len;\n\t\t\t\t\tconst T w = std::pow",
     "\t\t\t\t\tif(depth < T(0)) depth = T(0);\n\t\t\t\t\tconst T w = std::pow"),
    ("state step-up threshold 0.03", "if(mc < T(0.02) / T(st)) st = std::min(st + 1u, 8u);",
     "if(mc < T(0.03) / T(st)) st = std::min(st + 1u, 8u);"),
    ("all-out-of-range test at half the rays", "if(out_of_range >= n) {", "if(out_of_range >= n / 2) {"),
    # intersection, the watertight test's interval and the closest-hit tie break
    ("tmax exclusive", "return t >= r.tmin && t <= r.tmax;", "return t >= r.tmin && t < r.tmax;"),
    ("ties to the larger triangle id", "(t == best.t && id < best.tri)", "(t == best.t && id > best.tri)"),
    ("edge functions without the FP64 fallback", "\tif(U == 0.0f || V == 0.0f || W == 0.0f) {", "\tif(false) {"),
    # ray directions, sphericalFibonacci (src/shaders/irradiance.glsl:78-87), in oracle.hpp
    ("Fibonacci z without the half step", "T cos_t = T(1) - (T(2) * i + T(1)) * (T(1) / n);",
     "T cos_t = T(1) - (T(2) * i) * (T(1) / n);", "oracle.hpp"),
    # pbrMetallicRoughness (src/shaders/pbrMetallicRoughness.glsl:8-84)
    ("roughness not squared", "const T alpha = rough * rough;", "const T alpha = rough;"),
    ("reflectance at 90 degrees x50", "const T r90 = clampT(reflectance * T(25), T(0), T(1));",
     "const T r90 = clampT(reflectance * T(50), T(0), T(1));"),
    ("microfacet distribution clamp 8", "const T D = clampT(k * k * (T(1) / T(kPi)), T(0), T(4));",
     "const T D = clampT(k * k * (T(1) / T(kPi)), T(0), T(8));"),
    # border copy (src/shaders/probesCopyBorders.comp)
    ("left border not mirrored", "cp(0, j, 1, R - 1 - j);", "cp(0, j, 1, j);"),
    # probe init (src/shaders/probesInit.rgen:29-64) and selection (src/IrradianceProbes.cpp:396-424)
    ("inside-the-probe test threshold", "if(float(back) > 0.5f * float(n)) st = 0;", "if(float(back) > 0.25f * float(n)) st = 0;"),
    ("affecting geometry within two cells", "inside = inside && (dist < cell[k]);", "inside = inside && (dist < 2.0f * cell[k]);"),
    ("selection ignores the loop index", "((idx + *loop_index) % states[idx]) == 0", "(idx % states[idx]) == 0"),
    # atmospheric sky (src/shaders/sky.glsl:60-127)
    ("Mie asymmetry -0.98", "kSkyG = double(-0.990f)", "kSkyG = double(-0.980f)"),
    ("Rayleigh constant 0.0030", "kSkyKr = double(0.0025f)", "kSkyKr = double(0.0030f)"),
    # final gather (src/shaders/FinalGather.frag:38-77) and the ray-traced shadow pass
    # (src/shaders/directLight.rgen), in py.py and ddgi_oracle.cpp
    ("gather f0 0.04 instead of 0.004", "    f0 = 0.004\n", "    f0 = 0.04\n", "py.py"),
    ("gather without the emissive term", "    color += g[:, 12:15].astype(np.float64)\n", "\n", "py.py"),
    ("gather view direction reversed", "view = np.asarray(origin, np.float64)[None, :] - pos",
     "view = pos - np.asarray(origin, np.float64)[None, :]", "py.py"),
    ("shadow cone 0.04 rad", "2.0f * (nx - 0.5f) * 0.02f", "2.0f * (nx - 0.5f) * 0.04f"),
]


def main(only=()):
    tests = sorted(f for f in os.listdir(os.path.join(REPO, "tests")) if f.startswith("test_oracle_"))
    survived = []
    for name, old, new, *where in MUTATIONS:
        if only and name not in only:
            continue
        with tempfile.TemporaryDirectory() as d:
            for sub in ("oracle", "tests", "vulkanexp_amd", "include"):
                shutil.copytree(os.path.join(REPO, sub), os.path.join(d, sub),
                                ignore=shutil.ignore_patterns("_build", "_ref", "__pycache__", "*.so"))
            shutil.copy(os.path.join(REPO, "pytest.ini"), d)
            path = os.path.join(d, "oracle", where[0] if where else "ddgi_oracle.cpp")
            src = open(path).read()
            assert src.count(old) >= 1, f"{name}: the original text is missing"
            open(path, "w").write(src.replace(old, new, 1))  # the first occurrence: the forward code
            subprocess.run(["make", "-C", os.path.join(d, "oracle")], check=True, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL)
            r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider"] +
                               [os.path.join("tests", t) for t in tests], cwd=d, capture_output=True, text=True)
        caught = [l.split("::")[1].split(" ")[0] for l in r.stdout.splitlines() if l.startswith("FAILED")]
        print(f"{name:38s} {'caught by ' + ', '.join(caught) if r.returncode else 'SURVIVED'}", flush=True)
        if not r.returncode:
            survived.append(name)
    return survived


if __name__ == "__main__":
    sys.exit(1 if main(tuple(sys.argv[1:])) else 0)
