"""RIB scenes written for the tests: a textured still life in the manner of
render.rs's tests/fixtures/stilllife.rib (UDIM deck, graded label crate,
triplanar ball with fractal roughness, checker cloth, ramp band), built
from polygons, PointsPolygons with st and spheres, with synthetic .tex
files in place of the fixture's wood and label textures."""
import os

import numpy as np

import rtex
import scenes

STILL_LIFE = """
Format {w} {h} 1.0
Projection "perspective" "fov" [34]
Rotate -11 1 0 0
Translate 0 -2.6 0
WorldBegin
    Light "PxrDomeLight" "sky" "intensity" [0.6]
    LightSource "distantlight" 1 "from" [2 8 -4] "to" [0 0 0] "intensity" [1.5]
    AttributeBegin
        Attribute "identifier" "name" ["deck"]
        Pattern "PxrTexture" "wood" "filename" ["wood_<UDIM>.tex"]
        Bxdf "PxrSurface" "deck" "reference color diffuseColor" ["wood:resultRGB"]
            "diffuseRoughness" [0.3] "specularIor" [1.4] "specularRoughness" [0.35]
        PointsPolygons [4] [0 1 2 3]
            "P" [-9 0 2  9 0 2  9 0 20  -9 0 20]
            "st" [0 0  1.999 0  1.999 1.999  0 1.999]
    AttributeEnd
    AttributeBegin
        Attribute "identifier" "name" ["crate"]
        Pattern "PxrTexture" "poster" "filename" ["label.tex"]
        Pattern "PxrColorCorrect" "posterGraded" "reference color inputRGB" ["poster:resultRGB"]
            "gain" [1.05 1.0 0.92] "gamma" [1.1] "saturation" [1.25]
        Bxdf "PxrSurface" "crate" "reference color diffuseColor" ["posterGraded:resultRGB"]
            "specularIor" [1.45] "specularRoughness" [0.3]
        Translate -3.1 0.85 10
        Rotate 24 0 1 0
        PointsPolygons [4 4 4] [0 1 2 3  4 5 6 7  8 9 10 11]
            "P" [-0.85 -0.85 -0.85   0.85 -0.85 -0.85   0.85 0.85 -0.85  -0.85 0.85 -0.85
                 -0.85 -0.85  0.85  -0.85 -0.85 -0.85  -0.85 0.85 -0.85  -0.85 0.85  0.85
                 -0.85  0.85 -0.85   0.85  0.85 -0.85   0.85 0.85  0.85  -0.85 0.85  0.85]
            "st" [0 0  1 0  1 1  0 1  0 0  1 0  1 1  0 1  0 0  1 0  1 1  0 1]
    AttributeEnd
    AttributeBegin
        Attribute "identifier" "name" ["woodball"]
        Pattern "triplanar" "woodwrap" "filename" ["wood_1001.tex"] "frequency" [1.6]
        Pattern "PxrFractal" "rough" "frequency" [3.0] "layers" [4]
        Bxdf "PxrSurface" "woodball" "reference color diffuseColor" ["woodwrap:resultRGB"]
            "specularIor" [1.5] "reference float specularRoughness" ["rough:resultRGB"]
        Translate 3.1 1.0 11
        Sphere 1.0 -1.0 1.0 360
    AttributeEnd
    AttributeBegin
        Attribute "identifier" "name" ["cloth"]
        Pattern "PxrChecker" "check" "colorA" [0.75 0.12 0.10] "colorB" [0.92 0.88 0.82]
            "sScale" [10] "tScale" [10]
        Pattern "PxrMix" "checkSoft" "reference color color1" ["check:resultRGB"]
            "color2" [0.6 0.55 0.5] "mix" [0.12]
        Bxdf "PxrSurface" "cloth" "reference color diffuseColor" ["checkSoft:resultRGB"]
            "diffuseRoughness" [0.6] "specularIor" [1] "fuzzGain" [0.8] "fuzzColor" [0.95 0.9 0.85]
        Translate 0.1 0.001 12.5
        Rotate -14 0 1 0
        PointsPolygons [4] [0 1 2 3] "P" [-1.8 0 -1.8  1.8 0 -1.8  1.8 0 1.8  -1.8 0 1.8]
            "st" [0 0  1 0  1 1  0 1]
    AttributeEnd
    AttributeBegin
        Attribute "identifier" "name" ["band"]
        Pattern "PxrRamp" "fade" "axis" ["t"] "positions" [0 0.5 1]
            "colors" [0.06 0.22 0.28  0.1 0.4 0.45  0.75 0.82 0.8]
        Bxdf "PxrSurface" "band" "reference color diffuseColor" ["fade:resultRGB"]
            "specularIor" [1.5] "specularRoughness" [0.18]
        Translate -0.6 0 19
        PointsPolygons [4] [0 1 2 3] "P" [-6 0 0  6 0 0  6 3 1  -6 3 1] "st" [0 0  1 0  1 1  0 1]
    AttributeEnd
    AttributeBegin
        Attribute "identifier" "name" ["glassball"]
        Bxdf "PxrSurface" "glass" "diffuseGain" [0] "glassRefractionGain" [1] "glassIor" [1.5]
            "glassRoughness" [0.02] "refractionColor" [0.96 0.99 0.97]
        Translate 1.1 0.62 7.2
        Sphere 0.62 -0.62 0.62 360
    AttributeEnd
WorldEnd
"""


def write_still_life(directory, width=96, height=48):
    """Writes the RIB and its textures (wood tiles 1001, 1002, 1011, 1012
    and a label) into `directory`; returns the RIB's path."""
    for k, tile in enumerate((1001, 1002, 1011, 1012)):
        rtex.write_tex(os.path.join(directory, f"wood_{tile}.tex"),
                       scenes.smooth_texture(64, 64, seed=10 + k, cells=8) * np.float32(0.8))
    rtex.write_tex(os.path.join(directory, "label.tex"), scenes.stripes_texture(48, 48, 6) * np.float32(0.7) +
                   np.float32(0.15))
    path = os.path.join(directory, "still_life.rib")
    with open(path, "w") as f:
        f.write(STILL_LIFE.format(w=width, h=height))
    return path
