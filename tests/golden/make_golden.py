"""Write the golden vectors under tests/golden/ from the reference's own sources.

Runs oracle/_ref/chkn_ref (built by oracle/build_ref.sh from the reference tree; see
that script) in this container only.  Outputs, all data:

  flux_vectors.npz   seeded L1/L0/R0/R1 flow states + face frames -> reference face flux
                     (FVFace::calculate_convective_flux, face.cu:347-378), AUSMDV at x_order
                     1 and 2, then SBP-ASF (flux_calc 1) from a second seeded generator
  geom_vectors.npz   seeded warped hexahedra -> reference hex_cell_properties and
                     quad_properties (vector3.cu:196-245, :350-433)
  step_<name>.npz    small jobs marched by the reference driver: dt per step and the final
                     IO variables of every block, plus a digest of the job's inputs
  march_text.npz     a multi-block job in gzipped-text files marched to the end with
                     snapshots (times.data text, SHA-256 of every flow file written), then
                     restarted from time index 2 (the same after the restart)
  step_f32_<name>.npz  the same small jobs marched by the reference's FP32 build
                     (FLOAT_NUMBERS, oracle/_ref/chkn_ref*_f32), which reads gzipped-text
                     job files only: dt per step and the final IO variables as float32

Usage: python tests/golden/make_golden.py [step fixture name ...]   (default: all)
"""
import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "gdtk-rmhd_amd"))
REF = os.path.join(ROOT, "oracle", "_ref", "chkn_ref")
REF_AB = os.path.join(ROOT, "oracle", "_ref", "chkn_ref_ab")  # the reacting A->B gas build
REF_F32 = os.path.join(ROOT, "oracle", "_ref", "chkn_ref_f32")  # FLOAT_NUMBERS builds
REF_AB_F32 = os.path.join(ROOT, "oracle", "_ref", "chkn_ref_ab_f32")

from chkn import cases, prep  # noqa: E402
from chkn.job import BL_PROFILE_FILE, JobData, gas_flavour, read_flow_file, write_bl_profile  # noqa: E402

GAMMA, R = 1.4, 287.1
CV = R / (GAMMA - 1.0)


def random_states(rng, n, regime):
    """n flow states (rho, e, p, T, YB, a, vx, vy, vz) with p, T, a consistent."""
    T = rng.uniform(150.0, 3000.0, n)
    rho = 10.0 ** rng.uniform(-3.0, 1.0, n)
    a = np.sqrt(GAMMA * R * T)
    speed = {"sub": 0.6, "trans": 1.2, "super": 6.0}[regime]
    v = rng.normal(size=(n, 3)) * speed * a[:, None]
    p = rho * R * T
    e = T * CV
    return np.column_stack([rho, e, p, T, np.zeros(n), a, v])


def random_frames(rng, n):
    n_ = rng.normal(size=(n, 3))
    n_ /= np.linalg.norm(n_, axis=1)[:, None]
    t1 = np.cross(n_, rng.normal(size=(n, 3)))
    t1 /= np.linalg.norm(t1, axis=1)[:, None]
    t2 = np.cross(n_, t1)
    return np.column_stack([n_, t1, t2])


def flux_vectors(rng, nper=160):
    recs = []
    for regime in ("sub", "trans", "super"):
        for x_order in (1, 2):
            base = random_states(rng, nper, regime)
            # Stencil of four states near a common state: smooth, kinked and shocked lines.
            st = np.stack([base * (1.0 + 0.2 * rng.normal(size=base.shape) * rng.uniform(0, 1, (nper, 1)))
                           for _ in range(4)], axis=1)
            st[:, :, 3] = np.abs(st[:, :, 3])  # T
            st[:, :, 0] = np.abs(st[:, :, 0])  # rho
            st[:, :, 1] = st[:, :, 3] * CV
            st[:, :, 2] = st[:, :, 0] * R * st[:, :, 3]
            st[:, :, 5] = np.sqrt(GAMMA * R * st[:, :, 3])
            fr = random_frames(rng, nper)
            rec = np.column_stack([st.reshape(nper, 36), fr, np.full(nper, x_order), np.zeros(nper)])
            recs.append(rec)
    # Mirror pairs (wall-like): right states are left states with the normal velocity flipped.
    nm = 40
    base = random_states(rng, nm, "trans")
    fr = random_frames(rng, nm)
    st = np.repeat(base[:, None, :], 4, axis=1)
    for q in range(nm):
        nvec = fr[q, :3]
        for s in (2, 3):
            v = st[q, s, 6:9]
            st[q, s, 6:9] = v - 2.0 * np.dot(v, nvec) * nvec
    recs.append(np.column_stack([st.reshape(nm, 36), fr, np.full(nm, 2), np.zeros(nm)]))
    return np.concatenate(recs)


def sbp_flux_vectors(rng, nper=160):
    """Stencils of four nearby states for the SBP-ASF flux (flux_calc 1)."""
    recs = []
    for regime in ("sub", "trans", "super"):
        base = random_states(rng, nper, regime)
        st = np.stack([base * (1.0 + 0.1 * rng.normal(size=base.shape)) for _ in range(4)], axis=1)
        st[:, :, 3] = np.abs(st[:, :, 3])
        st[:, :, 0] = np.abs(st[:, :, 0])
        st[:, :, 1] = st[:, :, 3] * CV
        st[:, :, 2] = st[:, :, 0] * R * st[:, :, 3]
        st[:, :, 5] = np.sqrt(GAMMA * R * st[:, :, 3])
        st[:, :, 4] = np.abs(rng.normal(size=(nper, 4))) * (regime == "super")  # YB upwinding
        recs.append(np.column_stack([st.reshape(nper, 36), random_frames(rng, nper), np.full(nper, 2),
                                     np.ones(nper)]))
    return np.concatenate(recs)


def geom_vectors(rng, n=240):
    base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1],
                     [0, 1, 1]], dtype=float)
    out = []
    for q in range(n):
        scale = 10.0 ** rng.uniform(-3, 1, 3)
        A = rng.normal(size=(3, 3)) * 0.3 + np.eye(3)
        p = (base * scale) @ A.T + rng.normal(size=3)
        p += rng.normal(size=(8, 3)) * 0.08 * scale.min() * (q % 3)
        out.append(p.ravel())
    return np.array(out)


def run_ref(mode, rec):
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        np.ascontiguousarray(rec, dtype=float).tofile(fin)
        subprocess.run([REF, mode, fin, fout], check=True)
        return np.fromfile(fout, dtype=float)


def job_digest(job):
    f = job.flat()
    h = hashlib.sha256()
    for key in sorted(f):
        h.update(key.encode())
        h.update(np.ascontiguousarray(f[key]).tobytes())
    for key in ("x_order", "cfl_count", "max_step", "max_time", "dt_init", "viscous"):
        h.update(repr(getattr(job, key)).encode())
    return h.hexdigest()


STEP_JOBS = {
    # name: (builder, steps)
    "sod": (lambda: cases.sod_box(16, 8, 8, max_step=30), 30),
    "mixed": (lambda: cases.mixed_bc_array(max_step=25), 25),
    "ramp_cfl": (lambda: _cfl_limited_ramp(), 40),
    "xorder1": (lambda: _first_order(), 15),
    "tiny_n2": (lambda: _two_cell_block(), 20),
    "viscous": (lambda: cases.viscous_plate(max_step=30), 30),
    "viscous_x1": (lambda: cases.viscous_plate(max_step=20, x_order=1, seed=8), 20),
    "bl_plate": (lambda: cases.bl_plate(max_step=30), 30),
    "bl_plate_x1": (lambda: cases.bl_plate(max_step=20, x_order=1, seed=6), 20),
    "sbp": (lambda: cases.sbp_channel(max_step=25), 25),
    "sbp_viscous": (lambda: _viscous_sbp(), 20),
    "ab_reacting": (lambda: cases.ab_reacting(max_step=30), 30),
    "ab_inert": (lambda: cases.ab_reacting(max_step=20, seed=5, reacting=False), 20),
}


def _viscous_sbp():
    cases.sbp_channel(max_step=20, seed=5)
    prep.config.viscous = True
    return JobData.from_prep()


def _cfl_limited_ramp():
    cases.ramp(20, 6, 8, max_step=40)
    prep.config.dt_init = 2.0e-5  # dt falls to the CFL limit at step 10
    return JobData.from_prep()


def _first_order():
    cases.mixed_bc_array(max_step=15)
    prep.config.x_order = 1
    return JobData.from_prep()


def _two_cell_block():
    cases.ramp(2, 3, 2, max_step=20)
    return JobData.from_prep()


def step_fixture(name, builder, steps):
    built = builder()
    prof = built[1] if isinstance(built, tuple) else None
    job = built[0] if isinstance(built, tuple) else built
    ref = REF_AB if gas_flavour(job.gas_model) == "ab" else REF
    with tempfile.TemporaryDirectory() as d:
        prep.write_initial_files(True, d)
        if prof is not None:  # the solver reads the profile from its working directory
            write_bl_profile(os.path.join(d, BL_PROFILE_FILE), prof)
        dtlog = os.path.join(d, "dt.log")
        subprocess.run([ref, "run", d, "0", str(steps), dtlog], check=True, capture_output=True, cwd=d)
        dts = np.loadtxt(dtlog, ndmin=2)[:, 1]
        jd = JobData.from_job_dir(d, bl_profile=prof)
        fields = [read_flow_file(os.path.join(d, "flow", "t0001", "flow-%04d-%04d-%04d.bin" % (b.i, b.j, b.k)),
                                 b.ncells) for b in jd.blocks]
    np.savez_compressed(os.path.join(HERE, "step_%s.npz" % name), dt=dts, steps=steps,
                        digest=job_digest(jd), **{"io%d" % b: f for b, f in enumerate(fields)})


# Step jobs also marched by the FP32 build.
F32_STEP_JOBS = ["sod", "mixed", "ramp_cfl", "xorder1", "viscous", "bl_plate", "sbp", "ab_reacting"]


def f32_gas_model(job_dir):
    """Replace config.json's gas_model Cv by the FP32 build's own value.

    The FP32 build computes Cv = R/(g - 1.0) from g and R rounded to float (gas.cu:34-40);
    for the A->B gas that is 1434.9996..., which its 1e-4 gas-model check (config.cu:369-377)
    does not accept as the 1435 the prep writes.  Jobs for that build carry its own value."""
    import json
    path = os.path.join(job_dir, "config.json")
    with open(path) as f:
        cfg = json.load(f)
    gm = cfg["gas_model"]
    g, r = float(np.float32(gm["gamma"])), float(np.float32(gm["R"]))
    gm["Cv"] = float(np.float32(r / (g - 1.0)))
    with open(path, "w") as f:
        json.dump(cfg, f, indent=2)


def step_fixture_f32(name, builder, steps):
    built = builder()
    prof = built[1] if isinstance(built, tuple) else None
    job = built[0] if isinstance(built, tuple) else built
    ab = gas_flavour(job.gas_model) == "ab"
    ref = REF_AB_F32 if ab else REF_F32
    with tempfile.TemporaryDirectory() as d:
        prep.write_initial_files(False, d)  # block.cu:794-796: the FP32 build reads text only
        if ab:
            f32_gas_model(d)
        if prof is not None:
            write_bl_profile(os.path.join(d, BL_PROFILE_FILE), prof)
        dtlog = os.path.join(d, "dt.log")
        subprocess.run([ref, "run", d, "0", str(steps), dtlog, "textin"], check=True, capture_output=True,
                       cwd=d)
        dts = np.loadtxt(dtlog, ndmin=2)[:, 1].astype(np.float32)
        jd = JobData.from_job_dir(d, binary=False, bl_profile=prof)
        fields = [np.fromfile(os.path.join(d, "flow", "t0001", "flow-%04d-%04d-%04d.bin" % (b.i, b.j, b.k)),
                              dtype=np.float32, count=13 * b.ncells).reshape(13, b.ncells) for b in jd.blocks]
    np.savez_compressed(os.path.join(HERE, "step_f32_%s.npz" % name), dt=dts, steps=steps,
                        digest=job_digest(jd), **{"io%d" % b: f for b, f in enumerate(fields)})


def text_march_job():
    """mixed_bc_array for 40 steps with a snapshot every 1.5e-5 s (five snapshots)."""
    cases.mixed_bc_array(max_step=40)
    prep.config.dt_plot_list = [(0.0, 1.5e-5, 1.5e-5)]
    return JobData.from_prep()


def written_flow_digests(job_dir, first_tindx):
    """{relative path: sha256 of the decompressed text} of the flow files at time indices
    >= first_tindx."""
    import glob
    import gzip
    out = {}
    for path in sorted(glob.glob(os.path.join(job_dir, "flow", "t*", "*.gz"))):
        tindx = int(os.path.basename(os.path.dirname(path))[1:])
        if tindx >= first_tindx:
            with gzip.open(path, "rb") as f:
                out[os.path.relpath(path, job_dir)] = hashlib.sha256(f.read()).hexdigest()
    return out


def text_march_fixture():
    text_march_job()
    with tempfile.TemporaryDirectory() as d:
        prep.write_initial_files(False, d)
        ref = os.path.abspath(REF)
        subprocess.run([ref, "run", d, "0", "-1", "-", "text"], check=True, capture_output=True, cwd=d)
        with open(os.path.join(d, "times.data")) as f:
            times1 = f.read()
        dig1 = written_flow_digests(d, 1)
        subprocess.run([ref, "run", d, "2", "-1", "-", "text"], check=True, capture_output=True, cwd=d)
        with open(os.path.join(d, "times.data")) as f:
            times2 = f.read()
        dig2 = written_flow_digests(d, 3)
    np.savez_compressed(os.path.join(HERE, "march_text.npz"), times=times1, times_restart=times2,
                        names=np.array(sorted(dig1)), digests=np.array([dig1[k] for k in sorted(dig1)]),
                        names_restart=np.array(sorted(dig2)),
                        digests_restart=np.array([dig2[k] for k in sorted(dig2)]))


def main(names):
    if not os.path.exists(REF):
        sys.exit("build oracle/_ref first (oracle/build_ref.sh)")
    if names:  # only the named fixtures
        for name in names:
            if name == "march_text":
                text_march_fixture()
            elif name.startswith("f32_"):
                step_fixture_f32(name[4:], *STEP_JOBS[name[4:]])
            else:
                step_fixture(name, *STEP_JOBS[name])
        return
    rng = np.random.default_rng(20240601)
    rec = np.concatenate([flux_vectors(rng), sbp_flux_vectors(np.random.default_rng(20240607))])
    F = run_ref("flux", rec).reshape(-1, 6)
    np.savez_compressed(os.path.join(HERE, "flux_vectors.npz"), rec=rec, F=F)
    g = geom_vectors(rng)
    G = run_ref("geom", g).reshape(-1, 20)
    np.savez_compressed(os.path.join(HERE, "geom_vectors.npz"), p=g, out=G)
    for name, (builder, steps) in STEP_JOBS.items():
        step_fixture(name, builder, steps)
    for name in F32_STEP_JOBS:
        step_fixture_f32(name, *STEP_JOBS[name])
    text_march_fixture()
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1:])
