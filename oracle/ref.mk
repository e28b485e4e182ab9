# Builds the reference T-FSM miners, unmodified, from their own sources under
# REF (default /root/reference), into oracle/_ref/.  The products are test and
# baseline infrastructure only (golden pattern sets; the timed CPU baseline).
#
#   make -f oracle/ref.mk            # from the repository root
#
# -march=x86-64-v3 instead of the reference's -march=native: the AVX2 kernels
# in intersection/computesetintersection.cpp need AVX2, and the binary must
# also run on other x86-64 hosts than the one it was built on.
REF     ?= /root/reference
OUT     ?= oracle/_ref
CXX     ?= g++
CXXFLAGS = -std=c++14 -O3 -march=x86-64-v3 -fopenmp -w

all: $(OUT)/run_parallel $(OUT)/run_serial $(OUT)/run_fraction $(OUT)/ref_gmatch_dump $(OUT)/ref_fraction_dump

$(OUT)/parallel_intersection.o: $(REF)/parallel/intersection/computesetintersection.cpp
	@mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/parallel_run.o: $(REF)/parallel/run.cpp $(wildcard $(REF)/parallel/*.h)
	@mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/run_parallel: $(OUT)/parallel_run.o $(OUT)/parallel_intersection.o
	$(CXX) $(CXXFLAGS) $^ -o $@ -lpthread

$(OUT)/serial_intersection.o: $(REF)/serial/intersection/computesetintersection.cpp
	@mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/serial_run.o: $(REF)/serial/run.cpp $(wildcard $(REF)/serial/*.h)
	@mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/run_serial: $(OUT)/serial_run.o $(OUT)/serial_intersection.o
	$(CXX) $(CXXFLAGS) $^ -o $@

$(OUT)/fraction_intersection.o: $(REF)/fraction-score/intersection/computesetintersection.cpp
	@mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/fraction_run.o: $(REF)/fraction-score/run.cpp $(wildcard $(REF)/fraction-score/*.h)
	@mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -c $< -o $@

# the Fraction-Score build (support = min over pattern vertices of summed
# fractional scores, fraction-score/graph.h:895-922, comper.h:1082-1106)
$(OUT)/run_fraction: $(OUT)/fraction_run.o $(OUT)/fraction_intersection.o
	$(CXX) $(CXXFLAGS) $^ -o $@ -lpthread

# the fixture harness around the reference's matching engine
# (tests/golden/ref_gmatch_dump.cpp, our one translation unit, includes parallel/gmatch.h)
$(OUT)/ref_gmatch_dump: tests/golden/ref_gmatch_dump.cpp $(wildcard $(REF)/parallel/*.h) $(REF)/parallel/intersection/computesetintersection.cpp
	@mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -I$(REF)/parallel -o $@ tests/golden/ref_gmatch_dump.cpp $(REF)/parallel/intersection/computesetintersection.cpp -lpthread

# the fixture harness around the reference's Fraction-Score loader
# (tests/golden/ref_fraction_dump.cpp includes fraction-score/gmatch.h)
$(OUT)/ref_fraction_dump: tests/golden/ref_fraction_dump.cpp $(wildcard $(REF)/fraction-score/*.h) $(REF)/fraction-score/intersection/computesetintersection.cpp
	@mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -I$(REF)/fraction-score -o $@ tests/golden/ref_fraction_dump.cpp $(REF)/fraction-score/intersection/computesetintersection.cpp -lpthread

clean:
	rm -f $(OUT)/run_parallel $(OUT)/run_serial $(OUT)/run_fraction $(OUT)/ref_gmatch_dump $(OUT)/ref_fraction_dump $(OUT)/*.o

.PHONY: all clean
