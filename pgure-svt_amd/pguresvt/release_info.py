name = "pguresvt"
version = "0.6.4+mi355x.0.1.0"
description = "PGURE-SVT denoising of microscopy image sequences on AMD MI355X (gfx950)"
license = "GPLv3"
