# Builds the reference's own CPU solvers, from their sources where they lie,
# into oracle/_ref/ (git-ignored).  Used only to pin the oracle restatement
# and, optionally, as the "reference" CPU baseline.  Each solver is a single
# self-contained translation unit, compiled directly with gcc/g++.
#   make -f oracle/ref.mk            (needs the reference checkout at REF)
REF  ?= /root/reference
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT  := $(HERE)_ref
CXXFLAGS_REF := -O3 -std=c++11 -w
CFLAGS_REF   := -O3 -std=gnu99 -w

BINS := $(OUT)/idas_cpu_expand $(OUT)/idas_cpu_25_expand \
        $(OUT)/24puzzle_expand $(OUT)/24puzzle \
        $(OUT)/24puzzle_expand_with_pdb $(OUT)/24puzzle_with_pdb

all: $(BINS)

$(OUT):
	mkdir -p $(OUT)

$(OUT)/idas_cpu_expand: $(REF)/src_horie/idas_cpu_expand.c | $(OUT)
	gcc $(CFLAGS_REF) -o $@ $<
$(OUT)/idas_cpu_25_expand: $(REF)/src_horie/idas_cpu_25_expand.c | $(OUT)
	gcc $(CFLAGS_REF) -o $@ $<
$(OUT)/24puzzle_expand: $(REF)/src/24puzzle_expand.cc | $(OUT)
	g++ $(CXXFLAGS_REF) -o $@ $<
$(OUT)/24puzzle: $(REF)/src/24puzzle.cc | $(OUT)
	g++ $(CXXFLAGS_REF) -o $@ $<
$(OUT)/24puzzle_expand_with_pdb: $(REF)/src/24puzzle_expand_with_pdb.cpp | $(OUT)
	g++ $(CXXFLAGS_REF) -mcmodel=medium -o $@ $<
$(OUT)/24puzzle_with_pdb: $(REF)/src/24puzzle_with_pdb.cc | $(OUT)
	g++ $(CXXFLAGS_REF) -mcmodel=medium -o $@ $<

.PHONY: all
