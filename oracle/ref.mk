# ORACLE — test infrastructure only.
#
# Builds CUP-CFD's own fvm kernels, CupCfdSoAMesh and StructGen mesh source
# from the reference sources where they lie (REF, read-only) into
# oracle/_ref/cupcfd_ref_harness, driven by oracle/ref_harness.cpp. Outputs go
# to oracle/_ref/ only; nothing here is part of the product, and the GPU box
# never builds or runs it (it has no reference tree).
#
# Flags are the reference Release flags (CMakeLists.txt:83-90: -O2 -g, no
# -march, so x86-64 baseline arithmetic without FMA) with its C++17 standard
# (CMakeLists.txt:53). -fpermissive turns two ill-formed `return E;` in void
# functions (include/utility/interface/component/ArrayDrivers.ipp:342,351)
# into warnings, as the reference CI's GCC flags do (CMakeLists.txt:78); it
# does not change generated arithmetic. MPI is MPICH from /opt/conda (mpi.h,
# libmpi.so.12); the harness runs under mpiexec -n 2 (a one-rank mesh
# finalize fails in the reference, see ref_harness.cpp).
#
#   make -f oracle/ref.mk            # build oracle/_ref/cupcfd_ref_harness
REF ?= /root/reference
MPI ?= /opt/conda
CXX = g++
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_ref
OBJ := $(OUT)/obj

INCDIRS := $(shell find $(REF)/include -type d 2>/dev/null)
CXXFLAGS = -O2 -g -std=c++17 -fpermissive -w $(addprefix -I,$(INCDIRS)) -idirafter $(MPI)/include -DMPICH_SKIP_MPICXX
LDFLAGS = $(MPI)/lib/libmpi.so.12 -Wl,-rpath,$(MPI)/lib -static-libstdc++ -static-libgcc

SRCS := \
  src/comm/interface/component/Communicator.cpp \
  src/comm/interface/component/CustomMPIType.cpp \
  src/comm/interface/component/ExchangePattern.cpp \
  src/comm/implementation/component/MPIUtility.cpp \
  src/comm/implementation/component/ExchangePatternOneSidedNonBlocking.cpp \
  src/comm/implementation/component/ExchangePatternTwoSidedNonBlocking.cpp \
  src/comm/implementation/config/ExchangePatternConfig.cpp \
  src/data_structures/interface/component/AdjacencyList.cpp \
  src/data_structures/implementation/component/AdjacencyListCSR.cpp \
  src/data_structures/implementation/component/AdjacencyListVector.cpp \
  src/data_structures/implementation/component/DistributedAdjacencyList.cpp \
  src/error/Error.cpp \
  src/geometry/euclidean/implementation/component/EuclideanPlane3D.cpp \
  src/geometry/euclidean/implementation/component/EuclideanPoint.cpp \
  src/geometry/euclidean/implementation/component/EuclideanVector.cpp \
  src/geometry/euclidean/implementation/component/EuclideanVector2D.cpp \
  src/geometry/euclidean/implementation/component/EuclideanVector3D.cpp \
  src/geometry/euclidean/implementation/component/LineSegment3D.cpp \
  src/geometry/euclidean/implementation/component/Matrix.cpp \
  src/geometry/mesh/implementation/component/CupCfdSoAMesh.cpp \
  src/geometry/mesh/interface/component/UnstructuredMeshInterface.cpp \
  src/geometry/mesh/interface/component/UnstructuredMeshProperties.cpp \
  src/geometry/mesh/implementation/config/MeshConfig.cpp \
  src/geometry/mesh/implementation/config/MeshSourceStructGenConfig.cpp \
  src/geometry/mesh/interface/config/MeshSourceConfig.cpp \
  src/geometry/mesh/interface/source/MeshSource.cpp \
  src/geometry/mesh/implementation/source/MeshStructGenSource.cpp \
  src/geometry/shapes/implementation/component/Hexahedron.cpp \
  src/geometry/shapes/implementation/component/Pyramid.cpp \
  src/geometry/shapes/implementation/component/QuadPyramid.cpp \
  src/geometry/shapes/implementation/component/Quadrilateral3D.cpp \
  src/geometry/shapes/implementation/component/Tetrahedron.cpp \
  src/geometry/shapes/implementation/component/TriPrism.cpp \
  src/geometry/shapes/implementation/component/Triangle.cpp \
  src/geometry/shapes/implementation/component/Triangle3D.cpp \
  src/geometry/shapes/interface/component/Polygon.cpp \
  src/geometry/shapes/interface/component/Polygon3D.cpp \
  src/geometry/shapes/interface/component/Polyhedron.cpp \
  src/partitioner/interface/component/PartitionerInterface.cpp \
  src/partitioner/interface/config/PartitionerConfig.cpp \
  src/partitioner/implementation/component/PartitionerNaive.cpp \
  src/partitioner/implementation/config/PartitionerNaiveConfig.cpp

OBJS := $(OBJ)/ref_harness.o $(patsubst %.cpp,$(OBJ)/%.o,$(subst /,__,$(SRCS)))

$(OUT)/cupcfd_ref_harness: $(OBJS)
	$(CXX) -o $@ $(OBJS) $(LDFLAGS)

$(OBJ)/ref_harness.o: $(HERE)ref_harness.cpp
	@mkdir -p $(OBJ)
	$(CXX) $(CXXFLAGS) -c -o $@ $<

$(OBJ)/%.o:
	@mkdir -p $(OBJ)
	$(CXX) $(CXXFLAGS) -c -o $@ $(REF)/$(subst __,/,$*).cpp

clean:
	rm -rf $(OUT)

.PHONY: clean
