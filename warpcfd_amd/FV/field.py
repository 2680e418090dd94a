"""warp_cfd/FV/field.py:1-4."""


class Field:
    def __init__(self, name, index=-1):
        self.name = name
        self.index = index
