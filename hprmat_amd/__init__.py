"""hprmat_amd: MI355X-native dense complex LU + solve path of HPRMAT.

* ``hprmat_amd.solver``  - torch binding of libhprmat_hip (LU handles, solve)
* ``hprmat_amd.rmatrix`` - Lagrange-mesh R-matrix front end (mesh, assembly,
                           solver_type dispatch, R and S)
* ``hprmat_amd.synth``   - seeded synthetic systems of the benchmark workloads
* ``hprmat_amd.build``   - in-tree hipcc build of libhprmat_hip.so (gfx950)
"""
__version__ = '0.1.0'
