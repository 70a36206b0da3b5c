import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "smmp-svn_amd"))
sys.path.insert(0, os.path.join(REPO, "tests"))

GOLDEN = os.path.join(REPO, "tests", "golden")
LIBDIR = os.path.join(GOLDEN, "SMMP") + "/"
TOPO_TAGS = ["enkefa", "c2_20mer", "c5_36mer", "len10", "len60", "glyends", "pro_first", "allres"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X")


def golden(name: str):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def make_ctx(seq=None, tag=None, itysol=1, varfile=" ", tmpdir=None, sh2=0, seqfile=None):
    """Product context (ECEPP/3 as universe.py sets it up) for a sequence
    (a residue list, a fixture tag or a sequence file)."""
    from smmp_amd import _ffi, workloads
    c = _ffi.Context()
    c.view("epar_l.flex")[0] = 0
    c.view("epar_l.sh2")[0] = sh2
    c.view("epar_l.epsd")[0] = 0
    c.view("epar_l.ientyp")[0] = 0
    c.view("isolty.itysol")[0] = itysol
    c.init_energy(LIBDIR)
    if tag is not None:
        if tag == "enkefa":
            seqf = os.path.join(GOLDEN, "enkefa.seq")
        else:
            seq = list(golden(f"topo_{tag}.npz")["seq_in"])
    if seqfile is not None:
        seqf = seqfile
    if seq is not None:
        import tempfile
        d = tmpdir or tempfile.mkdtemp(prefix="smmp_t_")
        seqf = workloads.write_seq_file(os.path.join(d, "s.seq"), list(seq), "t")
    c.init_molecule(1, "nh2", "cooh", seqf, varfile)
    return c


def has_gpu() -> bool:
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


def write_multi_seq(path: str, chains) -> str:
    """A sequence file with one '#'-headed block per chain (redseq.f)."""
    with open(path, "w") as f:
        for k, ch in enumerate(chains):
            f.write(f"#  chain {k + 1}\n")
            f.write("   " + " ".join(r.capitalize() for r in ch) + "\n")
    return path


def make_multi_ctx(tag: str, itysol=None):
    """Context of a multi-chain fixture (tests/golden/multi_<tag>.npz) with
    the fixture's global parameters and box (the fixture's solvent unless
    itysol is given); returns (context, fixture)."""
    import tempfile
    g = golden(f"multi_{tag}.npz")
    chains = [str(c).split("-") for c in g["chains"]]
    d = tempfile.mkdtemp(prefix="smmp_m_")
    c = make_ctx(itysol=int(g["itysol_in"][0]) if itysol is None else itysol,
                 seqfile=write_multi_seq(os.path.join(d, "m.seq"), chains))
    gb = c.view("mol_r.gbpr")
    nml = len(chains)
    gb[:, :nml] = np.asarray(g["gbpr"]).reshape(nml, 6).T
    c.view("bounds.boxsize")[0] = float(g["boxsize_in"][0])
    return c, g
