"""The PML material extension (BoxMesh.material_map) against a brute-force
restatement of the reference's region-by-region rule
(step_3_PIC_Mesh_Partitioner/PIC_Heterogeneous_Material.F90:620-720 element
regions from the element centre, :744-850 node regions in three passes (walls
and mattress, then tubes, then corner boxes), :870-975 node material ids
(2 = touched by a PML element, 1 = on an interface plane of a wall or the
mattress), :988-1160 Node_Mat_Mapping, :1340-1349 the extension).  The
restatement loops over all node pairs like the reference, so the meshes are
small."""
import numpy as np
import pytest

from hirensi_amd.mesh import BoxMesh

TOL = 1.0e-6
WALLS, TUBES, BOXES = (2, 7, 8, 10, 13), (1, 3, 5, 9, 11, 12, 14, 16), (4, 6, 15, 17)
REGION = {(1, 1, 0): 1, (1, 0, 0): 2, (1, -1, 0): 3, (1, 1, -1): 4, (1, 0, -1): 5, (1, -1, -1): 6,
          (0, 1, 0): 7, (0, -1, 0): 8, (0, 1, -1): 9, (0, 0, -1): 10, (0, -1, -1): 11,
          (-1, 1, 0): 12, (-1, 0, 0): 13, (-1, -1, 0): 14, (-1, 1, -1): 15, (-1, 0, -1): 16,
          (-1, -1, -1): 17, (0, 0, 0): 18}


def reference_node_mat_mapping(mesh):
    xyz, inod = mesh.xyz(), mesh.inod()
    x01, x02 = mesh.xc[2 * mesh.rd_hi[0]], mesh.xc[2 * mesh.rd_lo[0]]
    x03, x04 = mesh.yc[2 * mesh.rd_hi[1]], mesh.yc[2 * mesh.rd_lo[1]]
    x06 = mesh.zc[2 * mesh.rd_lo[2]]
    nj = mesh.nnode
    ereg = np.zeros(mesh.nel, dtype=int)
    for e in range(mesh.nel):
        c = xyz[inod[e, 26]]                               # node 27: the element centre
        lx = 1 if c[0] >= x01 else (-1 if c[0] <= x02 else 0)
        ly = 1 if c[1] >= x03 else (-1 if c[1] <= x04 else 0)
        lz = -1 if c[2] <= x06 else 0
        ereg[e] = REGION[(lx, ly, lz)]
    nreg = np.full(nj, 18)
    for group in (WALLS, TUBES, BOXES):
        for e in range(mesh.nel):
            if ereg[e] in group:
                nreg[inod[e]] = ereg[e]
    mat_id = np.zeros(nj, dtype=int)
    for e in range(mesh.nel):
        if ereg[e] != 18:
            mat_id[inod[e]] = 2
    plane = {2: (0, x01), 13: (0, x02), 7: (1, x03), 8: (1, x04), 10: (2, x06)}
    for e in range(mesh.nel):
        if ereg[e] in plane:
            a, v = plane[ereg[e]]
            for n in inod[e]:
                if abs(xyz[n, a] - v) < TOL:
                    mat_id[n] = 1
    mapping = np.zeros(nj, dtype=np.int64) - 1
    indep = mat_id != 2
    mapping[indep] = np.flatnonzero(indep)
    # axes that must agree between a dependent node and its interface node, per region
    match = {1: (2,), 3: (2,), 12: (2,), 14: (2,), 4: (), 6: (), 15: (), 17: (),
             2: (1, 2), 13: (1, 2), 7: (0, 2), 8: (0, 2), 5: (1,), 16: (1,), 9: (0,), 11: (0,), 10: (0, 1)}
    for ij in np.flatnonzero(mat_id == 1):
        lreg = nreg[ij]
        if lreg not in match:
            continue
        for i in np.flatnonzero(nreg == lreg):
            if sum(abs(xyz[ij, a] - xyz[i, a]) for a in match[lreg]) < TOL:
                mapping[i] = ij
    return mapping, indep


@pytest.mark.parametrize("shape,npml", [((7, 6, 5), (2, 1, 1, 2, 2)), ((6, 6, 4), (1, 1, 1, 1, 1)),
                                        ((9, 8, 6), (3, 2, 2, 3, 2))])
def test_material_map_is_the_reference_extension(shape, npml):
    mesh = BoxMesh(*shape, npml=npml, h=(1.0, 1.25, 0.8))
    ref, indep = reference_node_mat_mapping(mesh)
    mm = mesh.material_map()
    dep = ~indep
    assert np.all(ref[dep] >= 0)                          # every PML node found its interface node
    assert np.array_equal(mm[dep], ref[dep])
    assert np.array_equal(mm[indep], np.flatnonzero(indep))
    # the extended material of a random field agrees
    rng = np.random.default_rng(1)
    lam = rng.uniform(50, 150, mesh.nnode)
    ext = lam.copy()
    ext[dep] = lam[ref[dep]]
    assert np.array_equal(lam[mm], ext)
