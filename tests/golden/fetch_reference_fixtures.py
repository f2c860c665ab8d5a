"""Copies the data fixtures of picca's correlation tests into tests/golden/.

Run once in a checkout that has the picca repository at /root/reference (or pass its
path).  Only data files are copied -- the delta inputs and the golden outputs that
``py/picca/tests/test_3_cor.py`` compares against -- never source code.

    python tests/golden/fetch_reference_fixtures.py [/path/to/picca]
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

FILES = {
    # inputs (py/picca/tests/data/test_delta)
    "py/picca/tests/data/test_delta/delta_attributes.fits.gz": "delta_attributes.fits.gz",
    "py/picca/tests/data/test_delta/Delta_LYA": "Delta_LYA",
    "py/picca/tests/data/test_delta/Delta_LYA_image": "Delta_LYA_image",
    # golden outputs (py/picca/tests/data/test_cor)
    "py/picca/tests/data/test_cor/cf.fits.gz": "cor/cf.fits.gz",
    "py/picca/tests/data/test_cor/cf_cross.fits.gz": "cor/cf_cross.fits.gz",
    "py/picca/tests/data/test_cor/cf_zcuts.fits.gz": "cor/cf_zcuts.fits.gz",
    "py/picca/tests/data/test_cor/cf_cross_zcuts.fits.gz": "cor/cf_cross_zcuts.fits.gz",
    "py/picca/tests/data/test_cor/cf_image.fits.gz": "cor/cf_image.fits.gz",
    "py/picca/tests/data/test_cor/cf_image_zcuts.fits.gz": "cor/cf_image_zcuts.fits.gz",
    "py/picca/tests/data/test_cor/cf_image_rebinned.fits.gz": "cor/cf_image_rebinned.fits.gz",
    "py/picca/tests/data/test_cor/cf_angl.fits.gz": "cor/cf_angl.fits.gz",
    "py/picca/tests/data/test_cor/dmat.fits.gz": "cor/dmat.fits.gz",
    "py/picca/tests/data/test_cor/dmat_cross.fits.gz": "cor/dmat_cross.fits.gz",
    "py/picca/tests/data/test_cor/dmat_zcuts.fits.gz": "cor/dmat_zcuts.fits.gz",
    "py/picca/tests/data/test_cor/dmat_cross_zcuts.fits.gz": "cor/dmat_cross_zcuts.fits.gz",
    # quasar catalogue and cross-correlation goldens
    "py/picca/tests/data/test_delta/cat.fits": "cat.fits",
    "py/picca/tests/data/test_cor/xcf.fits.gz": "cor/xcf.fits.gz",
    "py/picca/tests/data/test_cor/xcf_zcuts.fits.gz": "cor/xcf_zcuts.fits.gz",
    "py/picca/tests/data/test_cor/xdmat.fits.gz": "cor/xdmat.fits.gz",
    "py/picca/tests/data/test_cor/xdmat_zcuts.fits.gz": "cor/xdmat_zcuts.fits.gz",
    # metal distortion matrices, export, Wick covariance (and its 1-D correlation input)
    "py/picca/tests/data/test_cor/metal_dmat.fits.gz": "cor/metal_dmat.fits.gz",
    "py/picca/tests/data/test_cor/metal_dmat_cross.fits.gz": "cor/metal_dmat_cross.fits.gz",
    "py/picca/tests/data/test_cor/metal_xdmat.fits.gz": "cor/metal_xdmat.fits.gz",
    "py/picca/tests/data/test_cor/exported_cf.fits.gz": "cor/exported_cf.fits.gz",
    "py/picca/tests/data/test_cor/wick.fits.gz": "cor/wick.fits.gz",
    "py/picca/tests/data/test_cor/cf1d.fits.gz": "cor/cf1d.fits.gz",
}


def main(root="/root/reference"):
    for src, dst in FILES.items():
        src_path = os.path.join(root, src)
        dst_path = os.path.join(HERE, dst)
        os.makedirs(os.path.dirname(dst_path), exist_ok=True)
        if os.path.isdir(src_path):
            os.makedirs(dst_path, exist_ok=True)
            for name in sorted(os.listdir(src_path)):
                if name.endswith(".fits.gz"):
                    shutil.copyfile(os.path.join(src_path, name), os.path.join(dst_path, name))
        else:
            shutil.copyfile(src_path, dst_path)
        print(dst)


if __name__ == "__main__":
    main(*sys.argv[1:])
