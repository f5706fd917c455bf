"""Copies the image fixtures the reference's own tests hold into tests/golden/images/.

Data only (DDS files, PNG images from quicktex's tests/images/): the DDS files
pin the DDS reader and the block layout of a real file; Bun.png (512x622 RGBA)
is a real image with alpha for the DDS mip-chain encoder.

    python tests/golden/copy_images.py      # needs the reference tree
"""
import os
import shutil

SRC = '/root/reference/tests/images'
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'images')
FILES = ['bc1/3color.dds', 'bc1/3color.png', 'bc1/3color_black.dds', 'bc1/3color_black.png', 'bc1/greyscale.dds',
         'bc1/greyscale.png', 'bc4/6value.dds', 'bc4/6value.png', 'bc4/8value.dds', 'bc4/8value.png', 'Bun.png']

if __name__ == '__main__':
    for f in FILES:
        os.makedirs(os.path.dirname(os.path.join(DST, f)), exist_ok=True)
        shutil.copyfile(os.path.join(SRC, f), os.path.join(DST, f))
        print(f)
