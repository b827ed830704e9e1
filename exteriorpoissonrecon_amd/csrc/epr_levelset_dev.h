// Device-resident simplicial extraction (epr_levelset.hip): the triangulation
// of a list of grid cells and the zero sets of per-vertex values, with inputs
// and outputs left in device memory so that a pipeline can chain them.
#pragma once
#include <vector>

#include "epr_dev.h"

// a mesh handle of the C ABI: vertices (nV x vdim doubles) and simplices
// (nS x k vertex indices) in host memory
struct epr_mesh
{
	int vdim = 0 , k = 0;
	std::vector< double > v;
	std::vector< unsigned > s;
};

namespace epr
{
namespace ls
{
struct DevMesh
{
	DevBuf< double > V;        // nV x vdim
	DevBuf< unsigned > S;      // nS x k
	size_t nV = 0 , nS = 0;
	int vdim = 0 , k = 0;
};

// MarchingSimplices::RegularSubGridTriangulation( nCells , F , orient , false )
// (Misha/MarchingSimplices.h:622-706) of the device cell list dCells
// (nCells x dim), every coordinate of which is at most ext[d].
void TriangulateDev( int dim , const unsigned *dCells , size_t nCells , const unsigned ext[4] , bool orient , hipStream_t st , DevMesh &out );

// MarchingSimplices::LevelSet (Misha/MarchingSimplices.h:410-576) over a
// device mesh of Dim-simplices (dS: nS x (Dim+1)): codimension 2 (dValues:
// nV x 2 interleaved) or codimension 1 (dValues: nV), output vertices
// interpolating the rows of dVdata (nV x vdim).
template< int Dim >
void LevelSet2Dev( int vdim , size_t nV , const double *dVdata , size_t nS , const unsigned *dS , const double *dValues , const double iso[2] , hipStream_t st , DevMesh &out );
template< int Dim >
void LevelSet1Dev( int vdim , size_t nV , const double *dVdata , size_t nS , const unsigned *dS , const double *dValues , double iso , hipStream_t st , DevMesh &out );
} // namespace ls
} // namespace epr
