"""Writes the committed fixtures under tests/golden/ (run in the build
container; reads the reference's data files from /root/reference).

* ghia_re100.json    -- Re=100 columns of the Ghia, Ghia & Shin (1982) tables
                        shipped with the reference (examples/LDC/
                        u_velocity_results.txt, v_velocity_results.csv).
* cavity64_ref500.npz -- oracle, float32, 64x64 cavity, 500 SIMPLE
                        iterations with the reference's solver settings
                        (examples/LDC/LDC.py set-up), fields in reference
                        cell order.
* cavity24_converged.npz -- oracle float32 and float64, 24x24 cavity run to
                        the fixed point (700 iterations; float64 continuity
                        ~1e-12, Krylov atol 1e-13).
* cavity64_converged.npz -- oracle float64, the 64x64 cavity (LDC.py set-up)
                        run to continuity <= 1e-10 (Krylov atol 1e-13): the
                        fixed point the 500-iteration runs approach.
* mesh_3x3x2_jitter.npz -- connectivity / CSR tables of a jittered 3x3x2 hex
                        mesh from the brute-force pair loop.

    python tests/golden/make_golden.py [--skip-long | --cavity24 | --cavity64]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))
REF = '/root/reference/examples/LDC'


def ghia():
    u_rows = [l.split('\t') for l in open(os.path.join(REF, 'u_velocity_results.txt')).read().splitlines()]
    v_rows = [l.split(',') for l in open(os.path.join(REF, 'v_velocity_results.csv')).read().splitlines()]
    ui = u_rows[0].index('100')
    vi = v_rows[0].index('100')
    out = {'source': 'Ghia, Ghia & Shin (1982), Re=100, as shipped in examples/LDC of the reference',
           'u_on_x_0.5': [[float(r[0]), float(r[ui])] for r in u_rows[1:] if r and r[0]],
           'v_on_y_0.5': [[float(r[0]), float(r[vi])] for r in v_rows[1:] if r and r[0]]}
    json.dump(out, open(os.path.join(HERE, 'ghia_re100.json'), 'w'), indent=1)


def cavity(n, dtype, atol, iters, stop=None):
    from cases import oracle_case
    from oracle.simple import OracleSolver
    from warpcfd_amd.preprocess.grid import cavity_mesh
    nodes, cells = cavity_mesh(n)
    of = oracle_case(nodes, cells, 'cavity', dtype=dtype)
    s = OracleSolver(of, 0.7, 0.3, nu=0.01, atol=atol)
    hist = []
    for i in range(iters):
        info = s.outer_iteration()
        hist.append(info['continuity'])
        if i % 200 == 0:
            print(f'cavity{n} {np.dtype(dtype).name}: iteration {i}, continuity {hist[-1]:.3e}',
                  flush=True)
        if stop is not None and info['continuity'] < stop:
            break
    return of.cell_values.copy(), np.array(hist)


def mesh_tables():
    from oracle.mesh import OracleMesh
    from warpcfd_amd.preprocess.grid import structured_hex
    nodes, cells = structured_hex((3, 3, 2), (1.0, 1.0, 1.0), jitter=0.1, seed=11)
    m = OracleMesh(nodes, cells, pair_loop=True)
    r1, c1, _ = m.csr_pattern(1)
    r3, c3, _ = m.csr_pattern(3)
    np.savez_compressed(os.path.join(HERE, 'mesh_3x3x2_jitter.npz'), nodes=nodes, cells=cells,
                        adjacent_cells=m.adjacent_cells, cell_face_index=m.cell_face_index,
                        neighbors=m.neighbors, face_side=m.face_side, nnz_per_cell=m.nnz_per_cell,
                        face_offset_index=m.face_offset_index, cell_faces=m.cell_faces,
                        csr1_rowptr=r1, csr1_cols=c1, csr3_rowptr=r3, csr3_cols=c3,
                        psi=m.norm_distance[:, 0], area=m.face_area)


def main():
    if '--cavity24' in sys.argv:
        return cavity24()
    if '--cavity64' in sys.argv:
        return cavity64()
    ghia()
    mesh_tables()
    if '--skip-long' not in sys.argv:
        cv, hist = cavity(64, np.float32, None, 500)
        np.savez_compressed(os.path.join(HERE, 'cavity64_ref500.npz'), cell_values=cv, continuity=hist)
        cavity24()
        cavity64()


def cavity64():
    """64x64 Re 100 cavity, float64 oracle to continuity <= 1e-10."""
    c64, h64 = cavity(64, np.float64, 1e-13, 20000, stop=1e-10)
    np.savez_compressed(os.path.join(HERE, 'cavity64_converged.npz'), f64=c64,
                        iters64=len(h64), cont64=h64[-1], history=h64[::10])


def cavity24():
    """24x24 Re 100 cavity run to the fixed point: float64 to continuity
    ~1e-12 (700 iterations, Krylov atol 1e-13); float32 the same 700
    iterations (Krylov atol 1e-9), which ends at its rounding floor."""
    c32, h32 = cavity(24, np.float32, 1e-9, 700)
    c64, h64 = cavity(24, np.float64, 1e-13, 700)
    np.savez_compressed(os.path.join(HERE, 'cavity24_converged.npz'), f32=c32, f64=c64,
                        iters32=len(h32), iters64=len(h64), cont32=h32[-1], cont64=h64[-1])


if __name__ == '__main__':
    main()
