# Builds the unmodified reference (morris821028/hw-radiosity) from its sources
# where they lie, into oracle/_ref/ only.  Used to pin the oracle; never shipped.
#
#   make -f oracle/ref.mk            (from the repository root)
#
# Flags are the reference Makefile's (Makefile:7) plus `-include vector`,
# because src/rad.cpp:387 uses std::vector without including <vector>.
# Two numeric variants are produced:
#   REF-FMA    the shipped flags: GCC contracts a*b+c into FMA on this host
#   REF-NOFMA  the same with -ffp-contract=off (source-order arithmetic)
# refdump links the reference's own translation units with oracle/refdump.cpp
# (a dump harness), renaming the reference's main() with -Dmain=radref_main.
# rad_oldfmt is the shipped build with the reference's own OLD_FORMAT switch
# (rad.cpp:150) defined on the command line: the per-vertex colour output.
# tri2json is the WebGL viewer's model converter (WebGL_demo/model-converter/main.cpp).

REF     ?= /root/reference
HERE    := $(patsubst %/,%,$(dir $(lastword $(MAKEFILE_LIST))))
OUT     := $(HERE)/_ref
CXX     := g++
BASE    := -O2 -fopenmp -std=c++11 -march=native -msse -msse2 -Wall -Wno-unused-result \
           -Wno-misleading-indentation -Wno-unused-function -I$(REF)/include -include vector
FMA_FL  := $(BASE)
NOFMA_FL:= $(BASE) -ffp-contract=off
SRCS    := $(REF)/src/shade.cpp $(REF)/src/raycast.cpp $(REF)/src/rad.cpp
HDRS    := $(wildcard $(REF)/include/*.h)

# rad_v3 is the reference's own variant with Delta-ff adaptive splitting live:
# its config switches (include/config.h) set on the command line, with
# PARALLEL (centroid 3-way splits, PartitionTriangleWithCenterAndStore) and
# without DONT_PARTITION; _CONFIG_H keeps config.h itself from defining them.
V3_FL   := -D_CONFIG_H -DADAPT -DROLLBACK -DPRE_PARTITION -DPARALLEL

.PHONY: all clean
all: $(OUT)/rad $(OUT)/rad_nofma $(OUT)/rad_oldfmt $(OUT)/rad_v3 $(OUT)/refdump $(OUT)/refdump_nofma $(OUT)/refdump_v3 \
     $(OUT)/tri2json

$(OUT):
	mkdir -p $(OUT)

$(OUT)/rad: $(SRCS) $(HDRS) | $(OUT)
	$(CXX) $(FMA_FL) $(SRCS) -o $@ -lm

$(OUT)/rad_nofma: $(SRCS) $(HDRS) | $(OUT)
	$(CXX) $(NOFMA_FL) $(SRCS) -o $@ -lm

$(OUT)/rad_oldfmt: $(SRCS) $(HDRS) | $(OUT)
	$(CXX) $(FMA_FL) -DOLD_FORMAT $(SRCS) -o $@ -lm

$(OUT)/rad_v3: $(SRCS) $(HDRS) | $(OUT)
	$(CXX) $(FMA_FL) $(V3_FL) $(SRCS) -o $@ -lm

$(OUT)/%.v3.o: $(REF)/src/%.cpp $(HDRS) | $(OUT)
	$(CXX) $(FMA_FL) $(V3_FL) -Dmain=radref_main -c $< -o $@

$(OUT)/refdump_v3: $(HERE)/refdump.cpp $(OUT)/shade.v3.o $(OUT)/raycast.v3.o $(OUT)/rad.v3.o
	$(CXX) $(FMA_FL) $(V3_FL) $^ -o $@ -lm

$(OUT)/%.fma.o: $(REF)/src/%.cpp $(HDRS) | $(OUT)
	$(CXX) $(FMA_FL) -Dmain=radref_main -c $< -o $@

$(OUT)/%.nofma.o: $(REF)/src/%.cpp $(HDRS) | $(OUT)
	$(CXX) $(NOFMA_FL) -Dmain=radref_main -c $< -o $@

$(OUT)/refdump: $(HERE)/refdump.cpp $(OUT)/shade.fma.o $(OUT)/raycast.fma.o $(OUT)/rad.fma.o
	$(CXX) $(FMA_FL) $^ -o $@ -lm

$(OUT)/refdump_nofma: $(HERE)/refdump.cpp $(OUT)/shade.nofma.o $(OUT)/raycast.nofma.o $(OUT)/rad.nofma.o
	$(CXX) $(NOFMA_FL) $^ -o $@ -lm

# the WebGL viewer's model converter, one source file (no build file ships with it)
$(OUT)/tri2json: $(REF)/WebGL_demo/model-converter/main.cpp | $(OUT)
	$(CXX) -O2 -std=c++11 -w $< -o $@

clean:
	rm -rf $(OUT)
