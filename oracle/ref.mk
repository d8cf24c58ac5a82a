# Builds the reference solver from its own sources, where they lie, into
# oracle/_ref/ (never committed).  Used only to write golden fixtures and as
# an optional CPU baseline.  Requires the MPICH 3.3.2 shipped in the image
# under /opt/conda (headers + libmpi); nothing is stubbed.
#
#   make -f oracle/ref.mk            (from the repository root)
#
REF      ?= /root/reference/src
OUT      := oracle/_ref
MPI_ROOT ?= /opt/conda
CXX      := g++
# -D_GLIBCXX_USE_CXX11_ABI=0: the deck parser writes through the buffer of a
# temporary std::string (core/macros_inlines.h:46-52); it only survives with
# the copy-on-write string ABI.
CXXFLAGS := -O3 -D_GLIBCXX_USE_CXX11_ABI=0 -I$(MPI_ROOT)/include -w
SRCS := core/main.cpp core/simulation.cpp core/input.cpp core/state.cpp \
        overset/overset.cpp metrics/metrics.cpp \
        solver/solver.cpp solver/temporal_discretization.cpp solver/source.cpp \
        solver/spatial_discretization.cpp solver/boundary_condition.cpp \
        io/io.cpp io/plot3d.cpp io/overture.cpp io/tecplot.cpp io/fluent.cpp \
        geometry/geometry.cpp parallel/parallel.cpp parallel/decomp.cpp \
        math/math.cpp math/matrix.cpp
OBJS := $(patsubst %.cpp,$(OUT)/obj/%.o,$(SRCS))

all: $(OUT)/camille

$(OUT)/obj/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/camille: $(OBJS)
	$(CXX) -o $@ $(OBJS) $(MPI_ROOT)/lib/libmpicxx.so $(MPI_ROOT)/lib/libmpi.so \
	    -Wl,-rpath,/usr/lib/x86_64-linux-gnu -Wl,-rpath,$(MPI_ROOT)/lib

clean:
	rm -rf $(OUT)

.PHONY: all clean
