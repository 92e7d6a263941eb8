# Build the reference codec (SuperOptimizer/volume-compressor) from its own
# sources, in place under /root/reference, into oracle/_ref/ only.
#
# Used as a CHECKER: tests compare the oracle restatement (and, through the
# golden fixtures, the HIP library) against these builds, and bench.py may time
# libvcref_*.so as the CPU baseline. Nothing in the product links against it.
#
#   make -f oracle/ref.mk            (from the repo root)
#
# Variants:
#   libvcref_fast_v3.so / _v4.so  the reference's own CMake flags
#                                  (-O3 -ffp-contract=fast -ffast-math
#                                  -funroll-all-loops, CMakeLists.txt:46) at an
#                                  explicit ISA level (x86-64-v3 = AVX2+FMA,
#                                  x86-64-v4 = AVX-512) so the .so runs on other
#                                  x86 hosts than the one that built it.
#   libvcref_strict.so             -O2 with strict IEEE float (no fast-math, no
#                                  contraction): the pinned quantizer order.
#   libvcref_base.so               the reference flags with no -march (x86-64
#                                  baseline, no FMA instructions available).
#   vc_repack                      tools/vc_repack.c linked to libvcref_strict.so.
#   libdsref_fast.so / _fast_v3.so / _strict.so  tools/downscale.h (the
#                                  exporter's within-cell downscale) through
#                                  oracle/ds_ref_shim.c, with the reference flags
#                                  (baseline ISA, and x86-64-v3 with FMA) and
#                                  with strict float.
REF      ?= /root/reference
OUT      ?= oracle/_ref
SRC       = $(REF)/src/vc/vc.c
CFLAGS_C  = -std=c2x -fPIC -shared -DVC_TESTHOOKS -Wl,-Bsymbolic -Wno-misleading-indentation
FASTF     = -O3 -ffp-contract=fast -ffast-math -funroll-all-loops -fopenmp
LIBS      = -lm -lpthread

all: $(OUT)/libvcref_fast_v3.so $(OUT)/libvcref_fast_v4.so $(OUT)/libvcref_strict.so \
     $(OUT)/libvcref_base.so $(OUT)/vc_repack $(OUT)/libdsref_fast.so $(OUT)/libdsref_fast_v3.so \
     $(OUT)/libdsref_strict.so

$(OUT):
	mkdir -p $(OUT)

$(OUT)/libvcref_fast_v3.so: $(SRC) | $(OUT)
	gcc $(CFLAGS_C) $(FASTF) -march=x86-64-v3 $(SRC) -o $@ $(LIBS)

$(OUT)/libvcref_fast_v4.so: $(SRC) | $(OUT)
	gcc $(CFLAGS_C) $(FASTF) -march=x86-64-v4 $(SRC) -o $@ $(LIBS)

$(OUT)/libvcref_base.so: $(SRC) | $(OUT)
	gcc $(CFLAGS_C) $(FASTF) $(SRC) -o $@ $(LIBS)

$(OUT)/libvcref_strict.so: $(SRC) | $(OUT)
	gcc $(CFLAGS_C) -O2 -ffp-contract=off $(SRC) -o $@ $(LIBS)

$(OUT)/vc_repack: $(REF)/tools/vc_repack.c $(OUT)/libvcref_strict.so | $(OUT)
	gcc -std=c2x -O2 $(REF)/tools/vc_repack.c -o $@ -L$(OUT) -l:libvcref_strict.so \
	    -Wl,-rpath,'$$ORIGIN' $(LIBS)

$(OUT)/libdsref_fast.so: oracle/ds_ref_shim.c $(REF)/tools/downscale.h | $(OUT)
	gcc -std=c2x -fPIC -shared $(FASTF) -I$(REF)/tools oracle/ds_ref_shim.c -o $@

$(OUT)/libdsref_fast_v3.so: oracle/ds_ref_shim.c $(REF)/tools/downscale.h | $(OUT)
	gcc -std=c2x -fPIC -shared $(FASTF) -march=x86-64-v3 -I$(REF)/tools oracle/ds_ref_shim.c -o $@

$(OUT)/libdsref_strict.so: oracle/ds_ref_shim.c $(REF)/tools/downscale.h | $(OUT)
	gcc -std=c2x -fPIC -shared -O2 -ffp-contract=off -I$(REF)/tools oracle/ds_ref_shim.c -o $@

clean:
	rm -f $(OUT)/libvcref_*.so $(OUT)/libdsref_*.so $(OUT)/vc_repack

.PHONY: all clean
