# Builds the reference's own CLIs from their sources where they lie, into
# oracle/_ref/.  Nothing is copied: gcc reads the .c files in place.
#
#   make -f oracle/ref.mk            (from the repository root)
#
# Flags follow the reference's SubMakefile/MakeDef.mk:14-23
# (-D_GNU_SOURCE -std=gnu99 -O2, -lm -pthread) with one change:
# -march=x86-64-v3 instead of -march=native, so that the binary built here
# also runs on the host CPU of another machine.  x86-64-v3 has FMA and
# SSE4.1, so the two numerics-relevant code-generation facts are kept:
# gcc contracts Logspace's `a + i*c` into an FMA and CalcDistance is DPPS.
#
# These binaries are test infrastructure (golden vectors, CPU baseline);
# the product never links or runs them.

REF      ?= /root/reference
OUT      ?= oracle/_ref
CC       := gcc
CFLAGS   := -D_GNU_SOURCE -std=gnu99 -O2 -march=x86-64-v3 -Wno-unused-function \
            -Wno-unused-result -w -I$(REF)/include
LIBS     := -lm -pthread

S := $(REF)/src
COMMON   := $(S)/GetOption/GetOption.c
GADGET   := $(S)/Gadget/BufHalo.c $(S)/Gadget/BufVector3D.c \
            $(S)/Gadget/GadId.c $(S)/Gadget/GadgetHeader.c
POOLS    := $(S)/objects/CasQueue/CasQueue.c $(S)/objects/ThreadPool/ThreadPool.c \
            $(S)/objects/ThreadPool/ThreadRequest.c \
            $(S)/objects/SameThreadPool/SameThreadPool.c
ASTR     := $(S)/objects/ArraySTR/ArraySTR.c $(S)/objects/ArraySTR/Build.c \
            $(S)/objects/ArraySTR/MultiRangeCount.c $(S)/objects/ArraySTR/Node.c \
            $(S)/objects/RTree/Mbr.c $(S)/objects/RTree/MultiRangeCountContext.c
RTREE    := $(S)/objects/RTree/RTree.c $(S)/objects/RTree/Build.c \
            $(S)/objects/RTree/Node.c $(S)/objects/RTree/MultiRangeCount.c \
            $(S)/objects/RTree/Mbr.c $(S)/objects/RTree/MultiRangeCountContext.c

BINS := $(OUT)/astr_rcount $(OUT)/rt_rcount \
        $(OUT)/uniform_decoded_particle_generator $(OUT)/uniform_halo_generator

all: $(BINS)

$(OUT):
	mkdir -p $(OUT)

$(OUT)/astr_rcount: | $(OUT)
	$(CC) $(CFLAGS) -o $@ $(S)/astr_rcount/main.c $(COMMON) $(GADGET) $(POOLS) $(ASTR) $(LIBS)

$(OUT)/rt_rcount: | $(OUT)
	$(CC) $(CFLAGS) -o $@ $(S)/rt_rcount/main.c $(COMMON) $(GADGET) $(POOLS) $(RTREE) $(LIBS)

$(OUT)/uniform_decoded_particle_generator: | $(OUT)
	$(CC) $(CFLAGS) -o $@ $(S)/uniform_decoded_particle_generator/main.c $(COMMON) \
	    $(S)/mt19937ar/mt19937ar.c $(LIBS)

$(OUT)/uniform_halo_generator: | $(OUT)
	$(CC) $(CFLAGS) -o $@ $(S)/uniform_halo_generator/main.c $(COMMON) \
	    $(S)/mt19937ar/mt19937ar.c $(LIBS)

clean:
	rm -f $(BINS)

.PHONY: all clean
