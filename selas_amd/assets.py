"""Selas's material files turned into the 22-float material records sgpu_scene_set_materials takes.

* import_material: ImportMaterial (Source/Core/BuildCommon/ImportMaterial.cpp:56-87), the
  per-material JSON files of Selas's content (Content/Materials/*.json): "shaderName", "baseColor"
  (three numbers, or a texture name: then the colour is 0.6 grey), "transmittanceColor", the
  fourteen scalar attributes with their defaults (ImportMaterial.cpp:17-51; a string names a
  texture and leaves the default), "alphaTesting", "invertDisplacement", "usesPtex".
* import_disney_materials: ImportDisneyMaterials (ImportMaterial.cpp:90-147), the Disney
  Moana-island material file (named elements, "type" solid or thin, an optional "colorMap" per-face
  texture; each element's "assignment" names get a copy of the material).
* build_material: BuildMaterial (Source/Core/BuildCommon/BuildModel.cpp:22-63): flags, the shader
  by case-insensitive name (DisneySolid when none matches), the transparent flag of a DisneySolid
  material with diffTrans or specTrans above 0.

Numbers are read as rapidjson reads them (Json::ReadFloat / ReadFloat3, JsonUtilities.cpp:121-180):
a JSON number rounded to float; a value of another type takes the default. The record's layout is
the surface block of include/selas_gpu.h. Two things of Selas's materials have no place in it and
are returned beside the record: the displacement attribute (the device path has no displacement)
and texture names (UV textures are not read; per-face textures are set with sgpu_scene_set_ptex).
DeltaTransparent materials (Selas's eDiracTransparent) are refused: the device path has no such
shader.
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field

import numpy as np

SHADER_NAMES = ("DisneyThin", "DisneySolid", "DeltaTransparent")
ATTRIBUTES = ("metallic", "specularTint", "roughness", "anisotropic", "sheen", "sheenTint", "clearcoat",
              "clearcoatGloss", "specTrans", "diffTrans", "flatness", "ior", "scatterDistance", "displacement")
DEFAULTS = (0.3, 0.0, 0.5, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.5, 0.0, 0.0)
# record slot of each attribute (displacement has none)
SLOTS = {"metallic": 10, "specularTint": 16, "roughness": 17, "anisotropic": 14, "sheen": 6, "sheenTint": 7,
         "clearcoat": 8, "clearcoatGloss": 9, "specTrans": 11, "diffTrans": 12, "flatness": 13, "ior": 19,
         "scatterDistance": 18}
# MaterialFlags (ModelResource.h:36-43)
TRANSPARENT, ALPHA_TESTED, DISPLACEMENT_ENABLED, INVERT_DISPLACEMENT, USES_PTEX = 1, 2, 4, 8, 16


@dataclass
class ImportedMaterial:
    """ImportedMaterialData (ModelBuildPipeline.h:20-33)."""
    shader_name: str = ""
    base_color: tuple = (0.6, 0.6, 0.6)
    base_color_texture: str = ""
    transmittance_color: tuple = (0.0, 0.0, 0.0)
    scalars: dict = field(default_factory=dict)
    scalar_textures: dict = field(default_factory=dict)
    alpha_tested: bool = False
    invert_displacement: bool = False
    uses_ptex: bool = False


def _f32(x) -> float:
    return float(np.float32(x))


def _read_float(el: dict, key: str, default: float) -> float:
    v = el.get(key)
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        return _f32(default)
    return _f32(v)


def _read_float3(el: dict, key: str, default) -> tuple:
    v = el.get(key)
    if isinstance(v, list) and len(v) >= 3 and all(isinstance(x, (int, float)) and not isinstance(x, bool)
                                                   for x in v[:3]):
        return tuple(_f32(x) for x in v[:3])
    return tuple(_f32(x) for x in default)


def _read_bool(el: dict, key: str, default: bool) -> bool:
    v = el.get(key)
    return v if isinstance(v, bool) else default


def _load(src):
    if isinstance(src, dict):
        return src
    with open(src) as f:
        return json.load(f)


def import_material(src) -> ImportedMaterial:
    """ImportMaterial: a material JSON file (path) or its parsed dict."""
    doc = _load(src)
    m = ImportedMaterial()
    name = doc.get("shaderName")
    m.shader_name = name if isinstance(name, str) else ""
    if isinstance(doc.get("baseColor"), str):
        m.base_color_texture = doc["baseColor"]
        m.base_color = (_f32(0.6),) * 3
    else:
        m.base_color = _read_float3(doc, "baseColor", (0.6, 0.6, 0.6))
    m.transmittance_color = _read_float3(doc, "transmittanceColor", (0.0, 0.0, 0.0))
    for attr, default in zip(ATTRIBUTES, DEFAULTS):
        if isinstance(doc.get(attr), str):
            m.scalar_textures[attr] = doc[attr]
            m.scalars[attr] = _f32(default)
        else:
            m.scalars[attr] = _read_float(doc, attr, default)
    m.alpha_tested = _read_bool(doc, "alphaTesting", False)
    m.invert_displacement = _read_bool(doc, "invertDisplacement", False)
    m.uses_ptex = _read_bool(doc, "usesPtex", False)
    return m


def import_disney_materials(src) -> list:
    """ImportDisneyMaterials: [(name, ImportedMaterial)] in file order, each element followed by
    copies under its "assignment" names."""
    doc = _load(src)
    out = []
    for key, el in doc.items():
        m = ImportedMaterial(transmittance_color=(1.0, 1.0, 1.0))
        m.base_color = _read_float3(el, "baseColor", (1.0, 1.0, 1.0))
        for attr, default in zip(ATTRIBUTES, DEFAULTS):
            m.scalars[attr] = _read_float(el, attr, default)
        m.shader_name = "DisneySolid" if el.get("type") == "solid" else "DisneyThin"
        if "colorMap" in el:
            if not isinstance(el["colorMap"], str):
                raise ValueError(f"failed to find colormap for material {key}")
            m.base_color_texture = el["colorMap"].replace("\\", "/")
            m.uses_ptex = len(m.base_color_texture) > 0
        out.append((key, m))
        for a in el.get("assignment", []):
            out.append((a, m))
    return out


def build_material(m: ImportedMaterial) -> dict:
    """BuildMaterial: {"record": 22 floats, "shader": index, "flags", "displacement",
    "textures": {attribute or "baseColor": name}}."""
    flags = 0
    if m.alpha_tested:
        flags |= ALPHA_TESTED
    if m.invert_displacement:
        flags |= INVERT_DISPLACEMENT
    if m.uses_ptex:
        flags |= USES_PTEX
    shader = 1
    for k, n in enumerate(SHADER_NAMES):
        if n.lower() == m.shader_name.lower():
            shader = k
    if shader == 2:
        raise ValueError("DeltaTransparent materials are not supported by the device path")
    if shader == 1 and (m.scalars["diffTrans"] > 0.0 or m.scalars["specTrans"] > 0.0):
        flags |= TRANSPARENT
    rec = np.zeros(22, np.float32)
    rec[0:3] = m.base_color
    rec[3:6] = m.transmittance_color
    for attr, slot in SLOTS.items():
        rec[slot] = m.scalars[attr]
    rec[20] = shader
    rec[21] = flags
    textures = dict(m.scalar_textures)
    if m.base_color_texture:
        textures["baseColor"] = m.base_color_texture
    return {"record": rec, "shader": shader, "flags": flags, "displacement": m.scalars["displacement"],
            "textures": textures}


def load_material(path) -> np.ndarray:
    """The 22-float record of a Selas material file."""
    return build_material(import_material(path))["record"]
