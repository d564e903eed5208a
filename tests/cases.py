"""Writes small ctrl/grid/input sets for the parity tests.

Every case uses the repository's mechanism and thermo files
(eu3d-gpu_amd/input/H2_JA919.txt, thermo.dat) and the same key/value file
formats as the reference (input/ctrl.txt, grid-*.txt, input-*.txt).
"""
import os

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "eu3d-gpu_amd")
INPUT = os.path.join(PKG, "input")

RSBI_INPUT = """P_pre       101325.0
T_pre       {T_pre}
M_wave      {M_wave}
NS          9
NR          19
"""

SBI_INPUT = """case            SBI
P_pre           101325.0
T_pre           300.0
M_wave          2.0
NS              9
NR              19
bubble_xc_frac  0.35
bubble_r_frac   {r_frac}
shock_gap_cells {gap}
tanh_cells      2
perturb_amp     0.01
seed            {seed}
"""

ZND_INPUT = """case            ZND
P_pre           101325.0
T_pre           300.0
M_wave          0.0
NS              9
NR              19
driver_frac     0.05
driver_P        3039750.0
driver_T        3000.0
"""


ODW_INPUT = """case            ODW
P_pre           10132.5
T_pre           300.0
M_wave          7.0
NS              9
NR              19
wedge_deg       25.0
phi             1.0
"""


def write_case(tmpdir, name, ni, nj, nk, extents, input_text, time_adv="EE", dt=3e-8, cfl=0.0,
               final_time=None, steps=1, reaction="Trapezoid-Adaptive", difference="MUSCL_2"):
    """Write ctrl-<name>.txt (+ grid, input) into tmpdir; return the ctrl path."""
    xa, xb, ya, yb, za, zb = extents
    os.makedirs(tmpdir, exist_ok=True)
    grid = os.path.join(tmpdir, f"grid-{name}.txt")
    inp = os.path.join(tmpdir, f"input-{name}.txt")
    ctrl = os.path.join(tmpdir, f"ctrl-{name}.txt")
    with open(grid, "w") as f:
        f.write(f"total_ni {ni}\nxa {xa!r}\nxb {xb!r}\ntotal_nj {nj}\nya {ya!r}\nyb {yb!r}\n"
                f"total_nk {nk}\nza {za!r}\nzb {zb!r}\nbc 2\n")
    with open(inp, "w") as f:
        f.write(input_text)
    if final_time is None:
        final_time = steps * dt if cfl == 0 else 1.0
    with open(ctrl, "w") as f:
        f.write(f"gridname {grid}\ninitialization {inp}\n"
                f"reaction-model {os.path.join(INPUT, 'H2_JA919.txt')}\n"
                f"Thermofile {os.path.join(INPUT, 'thermo.dat')}\n"
                f"Final-Time {final_time:.12g}\ncfl {cfl!r}\ndt {dt!r}\ncount 1\n"
                f"difference {difference}\ntime-adv {time_adv}\nreaction {reaction}\n"
                f"m_block_x 1\nm_block_y 1\nm_block_z 1\nnum_thread 1\nDLB_step 5\nDLB_tol 0.001\n")
    return ctrl


def rsbi(tmpdir, ni=17, nj=13, nk=4, steps=3, time_adv="EE", T_pre=1000.0, M_wave=3.0, dt=3e-8,
         reaction="Trapezoid-Adaptive", difference="MUSCL_2"):
    """The reference's runnable problem (3dRSBI) on a reduced grid, same extents."""
    return write_case(tmpdir, f"rsbi{ni}x{nj}x{nk}", ni, nj, nk, (0.0, 0.17, 0.0, 0.08, 0.0, 0.02),
                      RSBI_INPUT.format(T_pre=T_pre, M_wave=M_wave), time_adv=time_adv, dt=dt, steps=steps,
                      reaction=reaction, difference=difference)


def sbi(tmpdir, ni=32, nj=16, nk=16, h=1e-4, gap=3, seed=20240601, time_adv="TVD_RK3", cfl=0.5, r_frac=0.2,
        reaction="Trapezoid-Adaptive", difference="MUSCL_2"):
    """Spherical shock-bubble (3dSBI) on a reduced grid with spacing h
    (a large negative r_frac leaves no bubble gas: pure nitrogen)."""
    return write_case(tmpdir, f"sbi{ni}x{nj}x{nk}", ni, nj, nk,
                      (0.0, (ni - 1) * h, 0.0, (nj - 1) * h, 0.0, (nk - 1) * h),
                      SBI_INPUT.format(gap=gap, seed=seed, r_frac=r_frac), time_adv=time_adv, cfl=cfl, dt=0.0,
                      reaction=reaction, difference=difference)


def znd(tmpdir, ni=201, time_adv="TVD_RK3", cfl=0.5, length=0.01):
    """1-D ZND initiation (1dZND) on a reduced grid."""
    h = length / (ni - 1)
    return write_case(tmpdir, f"znd{ni}", ni, 3, 3, (0.0, length, 0.0, 2 * h, 0.0, 2 * h), ZND_INPUT,
                      time_adv=time_adv, cfl=cfl, dt=0.0)


def odw(tmpdir, ni=32, nj=16, nk=4, h=1e-4, time_adv="TVD_RK3", cfl=0.5):
    """Wedge-frame oblique detonation (ODW: inflow x- and y+, wall y-) on a
    reduced grid with spacing h: Mach 7 H2-air at 0.1 atm, 300 K, 25 deg."""
    return write_case(tmpdir, f"odw{ni}x{nj}x{nk}", ni, nj, nk,
                      (0.0, (ni - 1) * h, 0.0, (nj - 1) * h, 0.0, (nk - 1) * h),
                      ODW_INPUT, time_adv=time_adv, cfl=cfl, dt=0.0)
