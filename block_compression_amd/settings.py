"""Compression variants and encoder presets.

Mirrors the reference crate's public types:

* ``CompressionVariant`` -- ``BC1``..``BC5``, ``BC6H(settings)``,
  ``BC7(settings)`` (src/lib.rs:74-103), with ``bytes_per_row`` and
  ``blocks_byte_size`` (src/lib.rs:129-157).  Equality and hashing look only
  at the format, not the settings (src/lib.rs:105-120).
* ``BC6HSettings`` -- 5 x u32, repr(C) (src/settings.rs:9-15) and its presets
  (src/settings.rs:20-75).
* ``BC7Settings`` -- 19 x u32, repr(C) (src/settings.rs:80-90) and its presets
  (src/settings.rs:96-241).

``to_bytes()`` gives the exact 20- / 76-byte blob the C ABI takes.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import List, Optional, Union


@dataclass(frozen=True)
class BC6HSettings:
    slow_mode: int
    fast_mode: int
    refine_iterations_1p: int
    refine_iterations_2p: int
    fast_skip_threshold: int

    def to_bytes(self) -> bytes:
        return struct.pack(
            "<5I",
            self.slow_mode,
            self.fast_mode,
            self.refine_iterations_1p,
            self.refine_iterations_2p,
            self.fast_skip_threshold,
        )

    # settings.rs:20-28
    @staticmethod
    def very_fast() -> "BC6HSettings":
        return BC6HSettings(0, 1, 0, 0, 0)

    # settings.rs:31-39
    @staticmethod
    def fast() -> "BC6HSettings":
        return BC6HSettings(0, 1, 0, 1, 2)

    # settings.rs:42-50
    @staticmethod
    def basic() -> "BC6HSettings":
        return BC6HSettings(0, 0, 2, 2, 4)

    # settings.rs:53-61
    @staticmethod
    def slow() -> "BC6HSettings":
        return BC6HSettings(1, 0, 2, 2, 10)

    # settings.rs:64-72
    @staticmethod
    def very_slow() -> "BC6HSettings":
        return BC6HSettings(1, 0, 2, 2, 32)


@dataclass(frozen=True)
class BC7Settings:
    refine_iterations: tuple
    mode_selection: tuple
    skip_mode2: int
    fast_skip_threshold_mode1: int
    fast_skip_threshold_mode3: int
    fast_skip_threshold_mode7: int
    mode45_channel0: int
    refine_iterations_channel: int
    channels: int

    def to_bytes(self) -> bytes:
        return struct.pack(
            "<19I",
            *self.refine_iterations,
            *self.mode_selection,
            self.skip_mode2,
            self.fast_skip_threshold_mode1,
            self.fast_skip_threshold_mode3,
            self.fast_skip_threshold_mode7,
            self.mode45_channel0,
            self.refine_iterations_channel,
            self.channels,
        )

    # settings.rs:96-108
    @staticmethod
    def opaque_ultra_fast() -> "BC7Settings":
        return BC7Settings((2, 2, 2, 1, 2, 2, 1, 0), (0, 0, 0, 1), 1, 3, 1, 0, 0, 0, 3)

    # settings.rs:111-123
    @staticmethod
    def opaque_very_fast() -> "BC7Settings":
        return BC7Settings((2, 2, 2, 1, 2, 2, 1, 0), (0, 1, 0, 1), 1, 3, 1, 0, 0, 0, 3)

    # settings.rs:126-138
    @staticmethod
    def opaque_fast() -> "BC7Settings":
        return BC7Settings((2, 2, 2, 1, 2, 2, 2, 0), (0, 1, 0, 1), 1, 12, 4, 0, 0, 0, 3)

    # settings.rs:141-153
    @staticmethod
    def opaque_basic() -> "BC7Settings":
        return BC7Settings((2, 2, 2, 2, 2, 2, 2, 0), (1, 1, 1, 1), 1, 12, 8, 0, 0, 2, 3)

    # settings.rs:156-168
    @staticmethod
    def opaque_slow() -> "BC7Settings":
        return BC7Settings((4, 4, 4, 4, 4, 4, 4, 0), (1, 1, 1, 1), 0, 64, 64, 0, 0, 4, 3)

    # settings.rs:171-183
    @staticmethod
    def alpha_ultrafast() -> "BC7Settings":
        return BC7Settings((2, 1, 2, 1, 1, 1, 2, 2), (0, 0, 1, 1), 1, 0, 0, 4, 3, 1, 4)

    # settings.rs:186-198
    @staticmethod
    def alpha_very_fast() -> "BC7Settings":
        return BC7Settings((2, 1, 2, 1, 2, 2, 2, 2), (0, 1, 1, 1), 1, 0, 0, 4, 3, 2, 4)

    # settings.rs:201-213
    @staticmethod
    def alpha_fast() -> "BC7Settings":
        return BC7Settings((2, 1, 2, 1, 2, 2, 2, 2), (0, 1, 1, 1), 1, 4, 4, 8, 3, 2, 4)

    # settings.rs:216-228
    @staticmethod
    def alpha_basic() -> "BC7Settings":
        return BC7Settings((2, 2, 2, 2, 2, 2, 2, 2), (1, 1, 1, 1), 1, 12, 8, 8, 0, 2, 4)

    # settings.rs:231-243
    @staticmethod
    def alpha_slow() -> "BC7Settings":
        return BC7Settings((4, 4, 4, 4, 4, 4, 4, 4), (1, 1, 1, 1), 0, 64, 64, 64, 0, 4, 4)


BC7_PRESETS = {
    "opaque_ultra_fast": BC7Settings.opaque_ultra_fast,
    "opaque_very_fast": BC7Settings.opaque_very_fast,
    "opaque_fast": BC7Settings.opaque_fast,
    "opaque_basic": BC7Settings.opaque_basic,
    "opaque_slow": BC7Settings.opaque_slow,
    "alpha_ultrafast": BC7Settings.alpha_ultrafast,
    "alpha_very_fast": BC7Settings.alpha_very_fast,
    "alpha_fast": BC7Settings.alpha_fast,
    "alpha_basic": BC7Settings.alpha_basic,
    "alpha_slow": BC7Settings.alpha_slow,
}

BC6H_PRESETS = {
    "very_fast": BC6HSettings.very_fast,
    "fast": BC6HSettings.fast,
    "basic": BC6HSettings.basic,
    "slow": BC6HSettings.slow,
    "very_slow": BC6HSettings.very_slow,
}

# Format tags of the C ABI (include/bcn_mi355x.h), in lib.rs:74-103 order,
# then the signed BC6H extension (BC6H_SF16: not a variant of the reference).
BC1, BC2, BC3, BC4, BC5, BC6H, BC7, BC6H_SF16 = range(8)
_NAMES = ("bc1", "bc2", "bc3", "bc4", "bc5", "bc6h", "bc7", "bc6h_sf16")


class CompressionVariant:
    """A block format plus, for BC6H and BC7, its encoder settings."""

    __slots__ = ("tag", "settings")

    def __init__(self, tag: int, settings: Optional[Union[BC6HSettings, BC7Settings]] = None):
        if tag not in range(len(_NAMES)):
            raise ValueError(f"unknown variant tag {tag}")
        if tag in (BC6H, BC6H_SF16) and not isinstance(settings, BC6HSettings):
            raise TypeError("BC6H needs BC6HSettings")
        if tag == BC7 and not isinstance(settings, BC7Settings):
            raise TypeError("BC7 needs BC7Settings")
        if tag not in (BC6H, BC6H_SF16, BC7) and settings is not None:
            raise TypeError("only BC6H and BC7 take settings")
        self.tag = tag
        self.settings = settings

    # Constructors named like the Rust enum's variants.
    @staticmethod
    def BC1() -> "CompressionVariant":
        return CompressionVariant(BC1)

    @staticmethod
    def BC2() -> "CompressionVariant":
        return CompressionVariant(BC2)

    @staticmethod
    def BC3() -> "CompressionVariant":
        return CompressionVariant(BC3)

    @staticmethod
    def BC4() -> "CompressionVariant":
        return CompressionVariant(BC4)

    @staticmethod
    def BC5() -> "CompressionVariant":
        return CompressionVariant(BC5)

    @staticmethod
    def BC6H(settings: BC6HSettings) -> "CompressionVariant":
        return CompressionVariant(BC6H, settings)

    @staticmethod
    def BC7(settings: BC7Settings) -> "CompressionVariant":
        return CompressionVariant(BC7, settings)

    @staticmethod
    def BC6H_SF16(settings: BC6HSettings) -> "CompressionVariant":
        """Signed BC6H (DXGI BC6H_SF16): the BC6H search over signed half
        floats.  An extension: the reference encodes unsigned BC6H only."""
        return CompressionVariant(BC6H_SF16, settings)

    @property
    def is_bc6h(self) -> bool:
        return self.tag in (BC6H, BC6H_SF16)

    # lib.rs:105-120: discriminant-only equality and hash.
    def __eq__(self, other) -> bool:
        return isinstance(other, CompressionVariant) and other.tag == self.tag

    def __hash__(self) -> int:
        return hash(self.tag)

    def __repr__(self) -> str:
        if self.settings is None:
            return f"CompressionVariant.{_NAMES[self.tag].upper()}"
        return f"CompressionVariant.{_NAMES[self.tag].upper()}({self.settings})"

    @property
    def name(self) -> str:
        return _NAMES[self.tag]

    # lib.rs:146-157
    def block_byte_size(self) -> int:
        return 8 if self.tag in (BC1, BC4) else 16

    # lib.rs:129-132
    def bytes_per_row(self, width: int) -> int:
        return ((width + 3) // 4) * self.block_byte_size()

    # lib.rs:138-144
    def blocks_byte_size(self, width: int, height: int) -> int:
        return ((width + 3) // 4) * ((height + 3) // 4) * self.block_byte_size()

    def settings_bytes(self) -> bytes:
        return b"" if self.settings is None else self.settings.to_bytes()
