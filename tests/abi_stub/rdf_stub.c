/*
 * TEST INFRASTRUCTURE ONLY: stands in for Moldy's rdf.c accumulator
 * (rdf_ptr, src/rdf.c:59-63) so that the drop-in force_calc's RDF branch can
 * be exercised without linking Moldy.  Loaded with RTLD_GLOBAL before
 * libmoldyforce.so, whose weak reference to rdf_ptr then binds here.
 */
static float *rdf_buf;
static int rdf_n;

void stub_set_rdf(float *p, int size)
{
   rdf_buf = p;
   rdf_n = size;
}

void *rdf_ptr(int *size)
{
   *size = rdf_n;
   return rdf_buf;
}
