"""Refined patches on the host: the patch geometry, the composite nodes and their volumes,
the per-patch fields and the caller's patch tables (``solver/refinement.py``,
``solver/tracer.py:1141-1183,1264-1276``), each pinned by a closed form."""

import numpy as np
import pytest

from beamtracer_amd import refinement
from beamtracer_amd.mesh import IonOpticsMesh
from beamtracer_amd.tracer import HipBeamTracer, node_volumes

MM = 1e-3


def _mesh(geometry="cylindrical", shape=(41, 21), h=0.25 * MM):
    return IonOpticsMesh(shape=shape, origin=(0.0, 0.0), h=h, geometry=geometry)


def test_patch_geometry_widens_to_cells_of_the_level_below():
    mesh = _mesh()
    base, fine = refinement.patches_of(mesh, [refinement.Refinement(1, (2.1 * MM, 0.0), (4.9 * MM, 1.3 * MM))])
    assert base.mesh is mesh and base.box == ((0, 40), (0, 20)) and base.field_mesh is None
    # x: 2.1 mm is base cell 8.4 -> node 8 -> fine node 16; 4.9 mm is 19.6 -> 20 -> 40
    # y: 0 -> 0; 1.3 mm is 5.2 -> 6 -> fine node 12
    assert fine.level == 1 and fine.box == ((16, 40), (0, 12))
    assert fine.mesh.shape == (25, 13) and fine.mesh.h == 0.125 * MM
    np.testing.assert_allclose(fine.bounds, (2.0 * MM, 0.0, 5.0 * MM, 1.5 * MM), rtol=0, atol=1e-15)
    # the field mesh: two fine cells more on every side inside the domain (none below the axis)
    assert fine.field_mesh.shape == (29, 15)
    np.testing.assert_allclose(fine.field_mesh.origin, (1.75 * MM, 0.0), atol=1e-15)
    assert fine.inner == (slice(2, 27), slice(0, 13))
    # a box reaching past the mesh is clipped to it; one outside raises
    (_, edge) = refinement.patches_of(mesh, [refinement.Refinement(1, (9.0 * MM, 4.0 * MM), (12.0 * MM, 9.0 * MM))])
    assert edge.box == ((72, 80), (32, 40)) and edge.field_mesh.shape == (11, 11)
    with pytest.raises(ValueError):
        refinement.patches_of(mesh, [refinement.Refinement(1, (11.0 * MM, 0.0), (12.0 * MM, 1.0 * MM))])


@pytest.mark.parametrize("geometry", ["cylindrical", "planar"])
def test_composite_nodes_and_volumes(geometry):
    mesh = _mesh(geometry)
    boxes = [refinement.Refinement(1, (2.1 * MM, 0.0), (4.9 * MM, 1.3 * MM)),
             refinement.Refinement(2, (2.6 * MM, 0.0), (3.4 * MM, 0.6 * MM))]  # fmt: skip
    patches = refinement.patches_of(mesh, boxes)
    layout = refinement.layout_of(mesh, patches)
    # the union of the nodes: each finer patch adds its nodes off the coarser lattice
    l1, l2 = patches[1].mesh.shape, patches[2].mesh.shape
    shared1 = ((l1[0] + 1) // 2) * ((l1[1] + 1) // 2)
    shared2 = ((l2[0] + 1) // 2) * ((l2[1] + 1) // 2)
    assert layout.size == 41 * 21 + (l1[0] * l1[1] - shared1) + (l2[0] * l2[1] - shared2)
    # every patch node is the composite node at its own position
    for p in patches:
        xn, yn = p.mesh.nodes()
        np.testing.assert_allclose(layout.x[p.index], xn, rtol=0, atol=1e-15)
        np.testing.assert_allclose(layout.y[p.index], yn, rtol=0, atol=1e-15)
    # the leaf cells tile the domain: the volumes add up to it exactly
    length, height = 10 * MM, 5 * MM
    total = np.pi * height**2 * length if geometry == "cylindrical" else height * length
    assert abs(layout.volumes.sum() - total) <= 1e-12 * total
    assert np.all(layout.volumes > 0)
    # a base node whose cells are all leaves and not on a face owns what node_volumes gives it
    far = layout.on_base(layout.volumes)[30, 5:15]
    np.testing.assert_allclose(far, node_volumes(mesh)[30, 5:15], rtol=1e-12)


def test_nesting_and_overlap_are_checked():
    mesh = _mesh()
    outside = [refinement.Refinement(1, (1.0 * MM, 0.0), (2.0 * MM, 1.0 * MM)),
               refinement.Refinement(2, (6.0 * MM, 0.0), (7.0 * MM, 1.0 * MM))]  # fmt: skip
    with pytest.raises(ValueError, match="outside every box"):
        refinement.layout_of(mesh, refinement.patches_of(mesh, outside))
    overlap = [refinement.Refinement(1, (1.0 * MM, 0.0), (3.0 * MM, 1.0 * MM)),
               refinement.Refinement(1, (2.0 * MM, 0.0), (4.0 * MM, 1.0 * MM))]  # fmt: skip
    with pytest.raises(ValueError, match="overlap"):
        refinement.layout_of(mesh, refinement.patches_of(mesh, overlap))


def test_deposit_total_adds_shared_nodes():
    mesh = _mesh()
    patches = refinement.patches_of(mesh, [refinement.Refinement(1, (2.1 * MM, 0.0), (4.9 * MM, 1.3 * MM))])
    layout = refinement.layout_of(mesh, patches)
    base = np.zeros(mesh.shape)
    fine = np.zeros(patches[1].mesh.shape)
    base[20, 3] = 1.0  # x = 5 mm, r = 0.75 mm: the fine patch's node (24, 6)
    fine[24, 6] = 2.0
    fine[1, 1] = 4.0  # a fine node of its own
    total = layout.deposit_total([base, fine])
    assert total[patches[0].index[20, 3]] == 3.0 and patches[0].index[20, 3] == patches[1].index[24, 6]
    assert total[patches[1].index[1, 1]] == 4.0 and total.sum() == 7.0
    assert np.array_equal(layout.on_base(total)[20, 3], 3.0)
    with pytest.raises(ValueError):
        layout.from_base(base)
    plain = refinement.layout_of(mesh, refinement.patches_of(mesh, []))
    np.testing.assert_array_equal(plain.on_base(plain.from_base(base)), base)


def test_patch_fields_of_a_linear_potential_are_uniform():
    mesh = _mesh("planar")
    patches = refinement.patches_of(mesh, [refinement.Refinement(1, (2.1 * MM, 1.1 * MM), (4.9 * MM, 3.3 * MM)),
                                           refinement.Refinement(2, (2.6 * MM, 1.6 * MM), (3.4 * MM, 2.4 * MM))])  # fmt: skip
    layout = refinement.layout_of(mesh, patches)
    ex, ey = 3.0e5, -1.25e5
    field = layout.field(100.0 - ex * layout.x - ey * layout.y)
    assert len(field.fields) == 3 and field.fx is field.fields[0].fx and field.mesh is mesh
    for f, m in zip(field.fields, field.field_meshes()):
        # the margins take the coarser patches' bilinear potential: exact for a linear one
        assert f.fx.shape == (m.shape[0] + 1, m.shape[1]) and f.fy.shape == (m.shape[0], m.shape[1] + 1)
        np.testing.assert_allclose(f.fx, ex, rtol=1e-9)
        np.testing.assert_allclose(f.fy, ey, rtol=1e-9)
    which = field.patch_of(np.array([1.0, 2.2, 3.0]) * MM, np.array([2.0, 2.0, 2.0]) * MM)
    assert which.tolist() == [0, 1, 2]
    np.testing.assert_allclose(field.spacing(np.array([1.0, 2.2, 3.0]) * MM, np.full(3, 2.0 * MM)),
                               np.array([0.25, 0.125, 0.0625]) * MM)  # fmt: skip
    with pytest.raises(ValueError):
        layout.field(np.zeros(layout.size - 1))


def test_caller_packs_patches_finest_first_and_unpacks_the_charges():
    mesh = _mesh()
    patches = refinement.patches_of(mesh, [refinement.Refinement(1, (2.1 * MM, 0.0), (4.9 * MM, 1.3 * MM))])
    layout = refinement.layout_of(mesh, patches)
    field = layout.field(np.zeros(layout.size))
    tracer = HipBeamTracer(mesh, field)
    fxs, fys, patch_int, patch_real = tracer.patch_arrays()
    fm = patches[1].field_mesh
    nf = fm.shape[0] * fm.shape[1]
    assert patch_int.tolist() == [[29, 15, 0, 0, nf, 0],
                                  [41, 21, 30 * 15, 29 * 16, 41 * 21, nf]]  # fmt: skip
    np.testing.assert_allclose(patch_real[0], (*fm.logical_origin, fm.h, 2.0 * MM, 0.0, 5.0 * MM, 1.5 * MM), atol=1e-15)
    np.testing.assert_allclose(patch_real[1], (0.0, 0.0, mesh.h, 0.0, 0.0, 10 * MM, 5 * MM), atol=1e-15)
    assert fxs.size == 30 * 15 + 42 * 21 and fys.size == 29 * 16 + 41 * 22
    # one unit on every node of both slabs: the fine box's margins drop, shared nodes add
    charge = tracer.node_charges(np.ones(int(patch_int[:, 4].sum())))
    counts = np.zeros(layout.size)
    for p in patches:
        np.add.at(counts, p.index.ravel(), 1.0)
    np.testing.assert_array_equal(charge, counts)
    np.testing.assert_allclose(tracer.finalize(charge), counts / layout.volumes)
    with pytest.raises(ValueError, match="refined"):
        tracer.require_uniform("the adjoint")
    # a field of the mesh alone keeps the (nx, ny) charges
    plain = HipBeamTracer(mesh, field.fields[0])
    assert plain.node_charges(np.arange(41 * 21.0)).shape == (41, 21) and plain.layout is None

