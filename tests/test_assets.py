"""Selas material files (selas_amd/assets.py: ImportMaterial, ImportDisneyMaterials, BuildMaterial)
against their closed form: the reference's own material files (Content/Materials, copied as
fixtures into tests/golden/materials) give records whose slots hold the files' values, the
importer's defaults where a value is absent (ImportMaterial.cpp:36-51: metallic 0.3, roughness
0.5, ior 1.5, the rest 0; baseColor 0.6 grey when absent or a texture name; transmittanceColor 0),
the shader by name and the transparent flag of a DisneySolid material with specTrans or diffTrans
above 0 (BuildModel.cpp:41-63). The reference's importer cannot be built here (it needs
AssetFileUtils.cpp, which does not compile on Linux)."""
import os

import numpy as np
import pytest

from selas_amd import assets as A

MATS = os.path.join(os.path.dirname(__file__), "golden", "materials")


def rec(base, trans=(0, 0, 0), sheen=0, sheen_tint=0, clearcoat=0, clearcoat_gloss=0, metallic=0.3, spec_trans=0,
        diff_trans=0, flatness=0, anisotropic=0, specular_tint=0, roughness=0.5, scatter=0, ior=1.5, shader=1, flags=0):
    r = np.zeros(22, np.float32)
    r[0:3] = base
    r[3:6] = trans
    r[6:15] = [sheen, sheen_tint, clearcoat, clearcoat_gloss, metallic, spec_trans, diff_trans, flatness, anisotropic]
    r[16:22] = [specular_tint, roughness, scatter, ior, shader, flags]
    return r


EXPECTED = {
    "Copper.json": rec((0.955008, 0.637427, 0.538163), ior=1.5, metallic=1.0, specular_tint=0.3, roughness=0.01,
                       clearcoat_gloss=1.0),
    "BSDF.json": rec((0.9, 1.0, 0.9), trans=(0.4519, 1.0, 0.7054), ior=1.7, metallic=0.0, roughness=0.01,
                     spec_trans=1.0, scatter=15.0, flags=A.TRANSPARENT),
    "Default.json": rec((0.6, 0.6, 0.6)),
    "DiffuseTransmissionThin.json": rec((0.9, 0.62, 0.9), ior=1.5, metallic=0.0, roughness=0.4, clearcoat_gloss=1.0,
                                        diff_trans=1.0, shader=0),
    "Paper.json": rec((0.6, 0.6, 0.6), metallic=0.0, roughness=0.8),
    "RoughDiffuse.json": rec((0.82, 0.39, 0.37), ior=1.2, metallic=0.0, roughness=0.8, flatness=0.5),
}


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_reference_material_files(name):
    b = A.build_material(A.import_material(os.path.join(MATS, name)))
    np.testing.assert_array_equal(b["record"], EXPECTED[name])
    assert b["displacement"] == 0.0


def test_texture_names_and_flags():
    b = A.build_material(A.import_material(os.path.join(MATS, "Paper.json")))
    assert b["textures"] == {"baseColor": "Textures~Paper~paper02_col.jpg"}
    m = A.import_material({"shaderName": "disneythin", "roughness": "rough.png", "usesPtex": True,
                           "alphaTesting": True, "invertDisplacement": True, "displacement": 0.25, "ior": 2,
                           "specTrans": 1.0})
    b = A.build_material(m)
    assert b["shader"] == 0 and b["flags"] == A.USES_PTEX | A.ALPHA_TESTED | A.INVERT_DISPLACEMENT
    assert b["record"][17] == np.float32(0.5) and b["textures"] == {"roughness": "rough.png"}
    assert b["record"][19] == 2.0 and b["displacement"] == np.float32(0.25)
    # thin materials are never flagged transparent; an unknown shader name is DisneySolid
    assert A.build_material(A.import_material({"shaderName": "Nope", "diffTrans": 0.5}))["flags"] == A.TRANSPARENT
    with pytest.raises(ValueError):
        A.build_material(A.import_material({"shaderName": "DeltaTransparent"}))


def test_disney_material_file():
    doc = {
        "isBeach": {"type": "solid", "baseColor": [0.7, 0.6, 0.5], "roughness": 0.9, "colorMap": "textures\\sand",
                    "assignment": ["sand_a", "sand_b"]},
        "isLeaf": {"type": "thin", "diffTrans": 0.4, "assignment": []},
    }
    out = A.import_disney_materials(doc)
    assert [n for n, _ in out] == ["isBeach", "sand_a", "sand_b", "isLeaf"]
    beach = A.build_material(out[0][1])
    np.testing.assert_array_equal(beach["record"], rec((0.7, 0.6, 0.5), trans=(1, 1, 1), roughness=0.9,
                                                       flags=A.USES_PTEX))
    assert beach["textures"] == {"baseColor": "textures/sand"}
    leaf = A.build_material(out[3][1])
    np.testing.assert_array_equal(leaf["record"], rec((1, 1, 1), trans=(1, 1, 1), diff_trans=0.4, shader=0))
